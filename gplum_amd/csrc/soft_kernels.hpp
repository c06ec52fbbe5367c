// soft_kernels.hpp -- kernel templates behind soft.hpp (included by the .hip units
// that instantiate them).
//
// Coordinates: every relative coordinate d = x - x_ref (formed in FP64) is held as
// a pair of floats hi = (float)d, lo = (float)(d - hi).  The neighbour predicate
// uses rx_pikg = xj_hi - xi_hi, which is exactly the PIKG kernel's relative
// coordinate (gravity_kernel_epep.pikg:53-60), so the FP32 neighbour count stays
// bit-identical to the oracle; the force of a near pair uses rx = rx_pikg + (xj_lo -
// xi_lo), whose error is relative to |r_ij| rather than to |x_j - x_ref|.  Far EP-EP
// pairs (beyond 1.1 max r_search) and EP-SP pairs use rx_pikg itself: at their
// distances an ulp of |x - x_ref| is small against |r_ij|.
#pragma once
#include "soft.hpp"

namespace gplum {

// Blocks are dealt to the 8 XCDs round-robin; map block b so that XCD x receives a
// contiguous run of groups (neighbouring groups share most of their j-lists, so
// they then share one L2).  The grid is padded to a multiple of 8.
__device__ inline int xcd_group(int ngroups) {
    const int per = (int)((gridDim.x + 7) / 8);
    return (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
}
inline int xcd_grid(int ngroups) { return ((ngroups + 7) / 8) * 8; }

// One lane set per split: wave-uniform because ni_pad is a multiple of 64.
struct SplitInfo {
    int ni_pad, nsplit, ii, sp;
    bool in_split, valid;
    __device__ SplitInfo(int ni) {
        ni_pad = (ni + 63) & ~63;
        if (ni_pad == 0) ni_pad = 64;
        nsplit = SOFT_BLOCK / ni_pad;
        ii = (int)threadIdx.x % ni_pad;
        // ni_pad is a multiple of 64: a wave's lanes share their split, so the j loops
        // below run on scalar counters and broadcast LDS reads
        sp = __builtin_amdgcn_readfirstlane((int)threadIdx.x / ni_pad);
        in_split = sp < nsplit;
        valid = in_split && ii < ni;
    }
};

struct IState {                 // i-particle registers
    float xh, yh, zh, xl, yl, zl, ro2, rs2;
    int id, rank;
};

// FP32 partial sums are folded into the FP64 accumulators every SOFT_FOLD pairs per
// lane: with FP32 pair terms that largely cancel (a particle inside a disk), longer
// FP32 runs double the rms error of the sum.
constexpr int SOFT_FOLD = 32;

struct Acc32 {                  // per-fold FP32 partial sums
    float x = 0.f, y = 0.f, z = 0.f, p = 0.f;
};

__device__ inline void split(double d, float &hi, float &lo) {
    hi = (float)d;
    lo = (float)(d - (double)hi);
}

// 1/sqrt(r2) to about 0.5 ulp: v_rsq_f32 (1 ulp) and one Newton step whose residual
// 1 - r2 y^2 is formed exactly (r2*y split into h + eh with an FMA), so the step adds
// only its final rounding.  Per pair this halves the error of rinv^3 against the
// textbook step y (3 - r2 y^2) / 2.
__device__ inline float rsqrt_refined(float r2) {
    const float y = __builtin_amdgcn_rsqf(r2);
    const float h = r2 * y;
    const float eh = __builtin_fmaf(r2, y, -h);
    const float e = __builtin_fmaf(-eh, y, __builtin_fmaf(-h, y, 1.0f));
    return __builtin_fmaf(0.5f * y, e, y);
}

// 1/sqrt(r2) by v_rsq_f32 and one textbook Newton step y (3 - r2 y^2) / 2
// (gravity_kernel_epep.pikg:89-91): for the far pairs, whose summands are small beside the
// near pairs' and carry no neighbour decision.
__device__ inline float rsqrt_newton(float r2) {
    const float y = __builtin_amdgcn_rsqf(r2);
    return y * __builtin_fmaf(-0.5f * r2, y * y, 1.5f);
}

// Appends (idi, idj) for the lanes with nb set: one ballot, one atomic per wave.
// Called by all lanes of a wave together.
__device__ inline void emit_candidate(bool nb, int idi, int idj, const CandBuf &cand) {
    const unsigned long long mask = __ballot(nb);
    if (mask) {
        const int lane = __lane_id();
        const int first = __ffsll((long long)mask) - 1;
        unsigned int b0 = 0;
        if (lane == first) b0 = atomicAdd(cand.count, (unsigned)__popcll(mask));
        b0 = (unsigned)__shfl((int)b0, first);
        if (nb) {
            const unsigned int pos = b0 + (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
            if (pos < cand.cap) cand.pairs[cand.base + pos] = make_int2(idi, idj);
        }
    }
}

// EP-EP pair with neighbour predicate, r_out clamp and candidate emission.
template <bool EMITC>
__device__ inline void epep_near(const IState &I, bool valid, float4 pj, float4 aj, float4 lj, float eps2,
                                 Acc32 &a, int &ngb, int &rankf, int &idmax, int &idmin,
                                 const CandBuf &cand) {
    const float rxp = pj.x - I.xh, ryp = pj.y - I.yh, rzp = pj.z - I.zh;
    const float r2r = __builtin_fmaf(rzp, rzp, __builtin_fmaf(ryp, ryp, __builtin_fmaf(rxp, rxp, eps2)));
    const float rs2 = fmaxf(I.rs2, aj.y);
    const int idj = __float_as_int(aj.z), rkj = __float_as_int(aj.w);
    const bool nb = valid && (r2r < rs2) && (I.id != idj || I.rank != rkj);
    if (nb) {
        ngb++;
        const int d = I.rank - rkj;
        rankf += d * d;
        idmax = max(idmax, idj);
        idmin = min(idmin, idj);
    }
    if (EMITC) emit_candidate(nb, I.id, idj, cand);
    const float rx = rxp + (lj.x - I.xl), ry = ryp + (lj.y - I.yl), rz = rzp + (lj.z - I.zl);
    const float r2f = __builtin_fmaf(rz, rz, __builtin_fmaf(ry, ry, __builtin_fmaf(rx, rx, eps2)));
    const float r2 = fmaxf(r2f, fmaxf(I.ro2, aj.x));
    const float rinv = rsqrt_refined(r2);
    const float mr = pj.w * rinv;
    const float mr3 = (rinv * rinv) * mr;
    a.x = __builtin_fmaf(mr3, rx, a.x);
    a.y = __builtin_fmaf(mr3, ry, a.y);
    a.z = __builtin_fmaf(mr3, rz, a.z);
    a.p -= mr;
}

// EP-EP pair known to lie beyond 1.1 max(r_search) > r_out: no predicate, no clamp.  At
// this distance the rounded parts alone give the relative coordinate (their error, half
// an ulp of |x - x_ref| each, is small against |r_ij|), and 1/r takes the textbook step.
__device__ inline void epep_far(const IState &I, float4 pj, float eps2, Acc32 &a) {
    const float rx = pj.x - I.xh;
    const float ry = pj.y - I.yh;
    const float rz = pj.z - I.zh;
    const float r2 = __builtin_fmaf(rz, rz, __builtin_fmaf(ry, ry, __builtin_fmaf(rx, rx, eps2)));
    const float rinv = rsqrt_newton(r2);
    const float mr = pj.w * rinv;
    const float mr3 = (rinv * rinv) * mr;
    a.x = __builtin_fmaf(mr3, rx, a.x);
    a.y = __builtin_fmaf(mr3, ry, a.y);
    a.z = __builtin_fmaf(mr3, rz, a.z);
    a.p -= mr;
}

// EP-SP pair: a node opened by neither the opening angle nor the search radius, as far
// from the group as the far EP-EP pairs or farther (rounded parts, textbook step).
template <bool QUAD>
__device__ inline void epsp_pair(const IState &I, float4 pj, float4 qa, float4 qb, float eps2, Acc32 &a) {
    const float rx = pj.x - I.xh;
    const float ry = pj.y - I.yh;
    const float rz = pj.z - I.zh;
    const float r2 = __builtin_fmaf(rz, rz, __builtin_fmaf(ry, ry, __builtin_fmaf(rx, rx, eps2)));
    const float rinv = rsqrt_newton(r2);
    const float r2inv = rinv * rinv;
    if (QUAD) {
        const float r3inv = r2inv * rinv;
        const float r4inv = r2inv * r2inv;
        const float r5inv = r2inv * r3inv;
        const float qrx = __builtin_fmaf(qa.x, rx, __builtin_fmaf(qa.w, ry, qb.y * rz));
        const float qry = __builtin_fmaf(qa.y, ry, __builtin_fmaf(qb.x, rz, qa.w * rx));
        const float qrz = __builtin_fmaf(qa.z, rz, __builtin_fmaf(qb.y, rx, qb.x * ry));
        const float rqr = __builtin_fmaf(qrz, rz, __builtin_fmaf(qry, ry, __builtin_fmaf(qrx, rx, qb.z)));
        const float rqr4 = rqr * r4inv;
        const float meff = __builtin_fmaf(0.5f, rqr4, pj.w);
        const float meff3 = __builtin_fmaf(2.5f, rqr4, pj.w) * r3inv;
        a.p = __builtin_fmaf(-meff, rinv, a.p);
        a.x = __builtin_fmaf(meff3, rx, __builtin_fmaf(-r5inv, qrx, a.x));
        a.y = __builtin_fmaf(meff3, ry, __builtin_fmaf(-r5inv, qry, a.y));
        a.z = __builtin_fmaf(meff3, rz, __builtin_fmaf(-r5inv, qrz, a.z));
    } else {
        const float mr = pj.w * rinv;
        const float mr3 = mr * r2inv;
        a.x = __builtin_fmaf(mr3, rx, a.x);
        a.y = __builtin_fmaf(mr3, ry, a.y);
        a.z = __builtin_fmaf(mr3, rz, a.z);
        a.p -= mr;
    }
}

template <class IS, class JS, class SS, class OUT, int MODE>
__global__ __launch_bounds__(SOFT_BLOCK) void k_soft_f32(const SoftGroup *__restrict__ groups,
                                                         int ngroups, IS is, JS js, JS jf, SS ss,
                                                         OUT out, float eps2, CandBuf cand, int tile) {
    const int g = xcd_group(ngroups);
    if (g >= ngroups) return;
    const SoftGroup G = groups[g];
    constexpr bool EMITC = (MODE & EMIT) != 0;
    const bool seg = EMITC && cand.seg_off != nullptr;
    if (G.ni <= 0) {
        if (seg && threadIdx.x == 0) cand.seg_cnt[g] = 0;
        return;
    }

    __shared__ float4 s_a[SOFT_TILE];
    __shared__ float4 s_b[SOFT_TILE];
    __shared__ float4 s_c[SOFT_TILE];
    __shared__ float4 s_l[SOFT_TILE];
    __shared__ double s_red[4][SOFT_BLOCK];
    __shared__ int s_redi[4][SOFT_BLOCK];
    __shared__ unsigned int s_cand;
    CandBuf cb = cand;
    if (seg) {
        if (threadIdx.x == 0) s_cand = 0u;   // visible after the first tile barrier
        cb.count = &s_cand;
        cb.base = (unsigned int)cand.seg_off[g];
        cb.cap = (unsigned int)cand.seg_cap[g];
    }

    double x0, y0, z0;
    {
        double ro, rs;
        int id, rk;
        is.load(G.i_ref, x0, y0, z0, ro, rs, id, rk);
    }
    const SplitInfo S(G.ni);
    const int t = (int)threadIdx.x;

    IState I{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, -1.f, -1, 0};
    if (S.valid) {
        double x, y, z, ro, rs;
        is.load(G.i_begin + S.ii, x, y, z, ro, rs, I.id, I.rank);
        split(x - x0, I.xh, I.xl);
        split(y - y0, I.yh, I.yl);
        split(z - z0, I.zh, I.zl);
        const float rof = (float)ro, rsf = (float)rs;
        I.ro2 = rof * rof;
        I.rs2 = (rsf * rsf) * 1.0201f;
    }
    double ax = 0., ay = 0., az = 0., pot = 0.;
    int ngb = 0, rankf = 0, idmax = ID_MAX_CLEAR, idmin = 2147483647;

    // NEAR: the near list (predicate, clamp, low parts); else the far list (rounded parts)
    auto stage_j = [&](const JS &src, int e, bool near) {
        double x, y, z, m, ro, rs;
        int id, rk;
        src.load(e, x, y, z, m, ro, rs, id, rk);
        float xh, xl, yh, yl, zh, zl;
        split(x - x0, xh, xl);
        split(y - y0, yh, yl);
        split(z - z0, zh, zl);
        s_a[t] = make_float4(xh, yh, zh, (float)m);
        if (near) {
            const float rof = (float)ro, rsf = (float)rs;
            s_l[t] = make_float4(xl, yl, zl, 0.f);
            s_b[t] = make_float4(rof * rof, (rsf * rsf) * 1.0201f, __int_as_float(id), __int_as_float(rk));
        }
    };

    if (MODE & EP) {
        // near list: predicate + clamp + candidates
        for (int base = 0; base < G.nj; base += tile) {
            const int cnt = min(tile, G.nj - base);
            __syncthreads();
            if (t < cnt) stage_j(js, G.j_begin + base + t, true);
            __syncthreads();
            if (S.in_split) {
                for (int k0 = S.sp; k0 < cnt; k0 += SOFT_FOLD * S.nsplit) {
                    Acc32 a;
                    const int k1 = min(cnt, k0 + SOFT_FOLD * S.nsplit);
                    for (int k = k0; k < k1; k += S.nsplit)
                        epep_near<EMITC>(I, S.valid, s_a[k], s_b[k], s_l[k], eps2, a, ngb, rankf, idmax,
                                         idmin, cb);
                    ax += a.x; ay += a.y; az += a.z; pot += a.p;
                }
            }
        }
        // far list (engine only): beyond 1.1 max(r_search) of the whole group
        for (int base = 0; base < ((MODE & NOFAR) ? 0 : G.njf); base += tile) {
            const int cnt = min(tile, G.njf - base);
            __syncthreads();
            if (t < cnt) stage_j(jf, G.jf_begin + base + t, false);
            __syncthreads();
            if (S.in_split) {
                for (int k0 = S.sp; k0 < cnt; k0 += SOFT_FOLD * S.nsplit) {
                    Acc32 a;
                    const int k1 = min(cnt, k0 + SOFT_FOLD * S.nsplit);
                    for (int k = k0; k < k1; k += S.nsplit) epep_far(I, s_a[k], eps2, a);
                    ax += a.x; ay += a.y; az += a.z; pot += a.p;
                }
            }
        }
    }

    if (MODE & (SPQ | SPM)) {
        for (int base = 0; base < G.ns; base += tile) {
            const int cnt = min(tile, G.ns - base);
            __syncthreads();
            if (t < cnt) {
                double m, x, y, z, q[6];
                ss.load(G.s_begin + base + t, m, x, y, z, q);
                s_a[t] = make_float4((float)(x - x0), (float)(y - y0), (float)(z - z0), (float)m);
                if (MODE & SPQ) {
                    // per-SPJ hoisted terms of gravity_kernel_epsp.pikg:75-82 (FP32)
                    const float qxx0 = (float)q[0], qyy0 = (float)q[1], qzz0 = (float)q[2];
                    const float qxy0 = (float)q[3], qxz0 = (float)q[4], qyz0 = (float)q[5];
                    const float tr = qxx0 + qyy0 + qzz0;
                    s_b[t] = make_float4(3.0f * qxx0 - tr, 3.0f * qyy0 - tr, 3.0f * qzz0 - tr, 3.0f * qxy0);
                    s_c[t] = make_float4(3.0f * qyz0, 3.0f * qxz0, -(eps2 * tr), 0.f);
                }
            }
            __syncthreads();
            if (S.in_split) {
                for (int k0 = S.sp; k0 < cnt; k0 += SOFT_FOLD * S.nsplit) {
                    Acc32 a;
                    const int k1 = min(cnt, k0 + SOFT_FOLD * S.nsplit);
                    for (int k = k0; k < k1; k += S.nsplit)
                        epsp_pair<(MODE & SPQ) != 0>(I, s_a[k], s_b[k], s_c[k], eps2, a);
                    ax += a.x; ay += a.y; az += a.z; pot += a.p;
                }
            }
        }
    }

    if (S.nsplit > 1) {
        __syncthreads();
        s_red[0][t] = ax; s_red[1][t] = ay; s_red[2][t] = az; s_red[3][t] = pot;
        s_redi[0][t] = ngb; s_redi[1][t] = rankf; s_redi[2][t] = idmax; s_redi[3][t] = idmin;
        __syncthreads();
        if (S.sp == 0 && S.valid) {
            for (int s = 1; s < S.nsplit; s++) {
                const int u = s * S.ni_pad + S.ii;
                ax += s_red[0][u]; ay += s_red[1][u]; az += s_red[2][u]; pot += s_red[3][u];
                ngb += s_redi[0][u]; rankf += s_redi[1][u];
                idmax = max(idmax, s_redi[2][u]); idmin = min(idmin, s_redi[3][u]);
            }
        }
    }
    if (S.sp == 0 && S.valid)
        out.write(G.i_begin + S.ii, ax, ay, az, pot, ngb, rankf, idmax, idmin, (MODE & EP) != 0);
    if (seg) {
        __syncthreads();
        if (threadIdx.x == 0) cand.seg_cnt[g] = (int32_t)s_cand;
    }
}

// FP64 twin: same structure, double throughout, r_inv = 1/sqrt(r2) (no Newton
// step); the near/far split is irrelevant in FP64 (both lists use the near loop).
// Without EMIT the neighbour predicate is evaluated in FP64 (the C-ABI twins).  With
// EMIT (the device system's FP64 runs) the neighbour count and the emitted candidates
// use the FP32 predicate of the FP32 kernel, so the neighbour decision that follows
// sees the same candidate set as in an FP32 run.
template <class IS, class JS, class SS, class OUT, int MODE>
__global__ __launch_bounds__(SOFT_BLOCK) void k_soft_f64(const SoftGroup *__restrict__ groups,
                                                         int ngroups, IS is, JS js, JS jf, SS ss,
                                                         OUT out, double eps2, CandBuf cand) {
    const int g = xcd_group(ngroups);
    if (g >= ngroups) return;
    const SoftGroup G = groups[g];
    if (G.ni <= 0) return;
    constexpr bool EMITC = (MODE & EMIT) != 0;
    const float eps2f = (float)eps2;

    constexpr int T64 = 128;
    __shared__ double s_d[12][T64];
    __shared__ int s_id[2][T64];
    __shared__ float s_h[4][T64];
    __shared__ double s_red[4][SOFT_BLOCK];
    __shared__ int s_redi[4][SOFT_BLOCK];

    double x0, y0, z0;
    {
        double ro, rs;
        int id, rk;
        is.load(G.i_ref, x0, y0, z0, ro, rs, id, rk);
    }
    const SplitInfo S(G.ni);
    const int t = (int)threadIdx.x;
    double xi = 0., yi = 0., zi = 0., ro2i = 0., rs2i = -1.;
    float xhi = 0.f, yhi = 0.f, zhi = 0.f, rs2fi = -1.f;
    int idi = -1, ranki = 0;
    if (S.valid) {
        double x, y, z, ro, rs;
        is.load(G.i_begin + S.ii, x, y, z, ro, rs, idi, ranki);
        xi = x - x0; yi = y - y0; zi = z - z0;
        ro2i = ro * ro;
        rs2i = (rs * rs) * 1.0201;
        xhi = (float)xi; yhi = (float)yi; zhi = (float)zi;
        const float rsf = (float)rs;
        rs2fi = (rsf * rsf) * 1.0201f;
    }
    double ax = 0., ay = 0., az = 0., pot = 0.;
    int ngb = 0, rankf = 0, idmax = ID_MAX_CLEAR, idmin = 2147483647;

    if (MODE & EP) {
        for (int pass = 0; pass < 2; pass++) {
            const JS &src = pass == 0 ? js : jf;
            const int nlist = pass == 0 ? G.nj : G.njf;
            const int lbeg = pass == 0 ? G.j_begin : G.jf_begin;
            for (int base = 0; base < nlist; base += T64) {
                const int cnt = min(T64, nlist - base);
                __syncthreads();
                if (t < cnt) {
                    double x, y, z, m, ro, rs;
                    int id, rk;
                    src.load(lbeg + base + t, x, y, z, m, ro, rs, id, rk);
                    s_d[0][t] = x - x0; s_d[1][t] = y - y0; s_d[2][t] = z - z0; s_d[3][t] = m;
                    s_d[4][t] = ro * ro; s_d[5][t] = (rs * rs) * 1.0201;
                    s_id[0][t] = id; s_id[1][t] = rk;
                    if (EMITC) {
                        const float rsf = (float)rs;
                        s_h[0][t] = (float)(x - x0); s_h[1][t] = (float)(y - y0); s_h[2][t] = (float)(z - z0);
                        s_h[3][t] = (rsf * rsf) * 1.0201f;
                    }
                }
                __syncthreads();
                if (S.in_split) {
                    for (int k = S.sp; k < cnt; k += S.nsplit) {
                        const double rx = s_d[0][k] - xi, ry = s_d[1][k] - yi, rz = s_d[2][k] - zi;
                        const double r2r = rx * rx + ry * ry + rz * rz + eps2;
                        const double rout2 = smax(ro2i, s_d[4][k]);
                        const double rs2 = smax(rs2i, s_d[5][k]);
                        const double r2 = smax(r2r, rout2);
                        const int idj = s_id[0][k], rkj = s_id[1][k];
                        bool within = r2r < rs2;
                        if (EMITC) {
                            const float rxp = s_h[0][k] - xhi, ryp = s_h[1][k] - yhi, rzp = s_h[2][k] - zhi;
                            const float r2f = __builtin_fmaf(rzp, rzp, __builtin_fmaf(ryp, ryp,
                                                                                     __builtin_fmaf(rxp, rxp, eps2f)));
                            within = r2f < fmaxf(rs2fi, s_h[3][k]);
                        }
                        const bool nb = S.valid && within && (idi != idj || ranki != rkj);
                        if (nb) {
                            ngb++;
                            const int d = ranki - rkj;
                            rankf += d * d;
                            idmax = max(idmax, idj);
                            idmin = min(idmin, idj);
                        }
                        if (EMITC) emit_candidate(nb, idi, idj, cand);
                        const double rinv = 1.0 / sqrt(r2);
                        const double mr = s_d[3][k] * rinv;
                        const double mr3 = rinv * rinv * mr;
                        ax += mr3 * rx; ay += mr3 * ry; az += mr3 * rz;
                        pot -= mr;
                    }
                }
            }
        }
    }
    if (MODE & (SPQ | SPM)) {
        for (int base = 0; base < G.ns; base += T64) {
            const int cnt = min(T64, G.ns - base);
            __syncthreads();
            if (t < cnt) {
                double m, x, y, z, q[6];
                ss.load(G.s_begin + base + t, m, x, y, z, q);
                s_d[0][t] = x - x0; s_d[1][t] = y - y0; s_d[2][t] = z - z0; s_d[3][t] = m;
                for (int k = 0; k < 6; k++) s_d[4 + k][t] = q[k];
            }
            __syncthreads();
            if (S.in_split) {
                for (int k = S.sp; k < cnt; k += S.nsplit) {
                    const double rx = s_d[0][k] - xi, ry = s_d[1][k] - yi, rz = s_d[2][k] - zi;
                    const double r2 = rx * rx + ry * ry + rz * rz + eps2;
                    const double rinv = 1.0 / sqrt(r2);
                    const double m = s_d[3][k];
                    if (MODE & SPQ) {
                        const double r2inv = rinv * rinv, r3inv = r2inv * rinv;
                        const double r4inv = r2inv * r2inv, r5inv = r2inv * r3inv;
                        const double tr = s_d[4][k] + s_d[5][k] + s_d[6][k];
                        const double qxx = 3. * s_d[4][k] - tr, qyy = 3. * s_d[5][k] - tr;
                        const double qzz = 3. * s_d[6][k] - tr, qxy = 3. * s_d[7][k];
                        const double qzx = 3. * s_d[8][k], qyz = 3. * s_d[9][k];
                        const double mtr = -(eps2 * tr);
                        const double qrx = qxx * rx + qxy * ry + qzx * rz;
                        const double qry = qyy * ry + qyz * rz + qxy * rx;
                        const double qrz = qzz * rz + qzx * rx + qyz * ry;
                        const double rqr = mtr + qrx * rx + qry * ry + qrz * rz;
                        const double rqr4 = rqr * r4inv;
                        const double meff = m + 0.5 * rqr4;
                        const double meff3 = (m + 2.5 * rqr4) * r3inv;
                        pot -= meff * rinv;
                        ax += -r5inv * qrx + meff3 * rx;
                        ay += -r5inv * qry + meff3 * ry;
                        az += -r5inv * qrz + meff3 * rz;
                    } else {
                        const double mr = m * rinv;
                        const double mr3 = mr * rinv * rinv;
                        ax += mr3 * rx; ay += mr3 * ry; az += mr3 * rz;
                        pot -= mr;
                    }
                }
            }
        }
    }
    if (S.nsplit > 1) {
        __syncthreads();
        s_red[0][t] = ax; s_red[1][t] = ay; s_red[2][t] = az; s_red[3][t] = pot;
        s_redi[0][t] = ngb; s_redi[1][t] = rankf; s_redi[2][t] = idmax; s_redi[3][t] = idmin;
        __syncthreads();
        if (S.sp == 0 && S.valid) {
            for (int s = 1; s < S.nsplit; s++) {
                const int u = s * S.ni_pad + S.ii;
                ax += s_red[0][u]; ay += s_red[1][u]; az += s_red[2][u]; pot += s_red[3][u];
                ngb += s_redi[0][u]; rankf += s_redi[1][u];
                idmax = max(idmax, s_redi[2][u]); idmin = min(idmin, s_redi[3][u]);
            }
        }
    }
    if (S.sp == 0 && S.valid)
        out.write(G.i_begin + S.ii, ax, ay, az, pot, ngb, rankf, idmax, idmin, (MODE & EP) != 0);
}

template <class IS, class JS, class SS, class OUT, int MODE>
void launch_soft_f32(const SoftGroup *d_groups, int ngroups, IS is, JS js, JS jf, SS ss, OUT out,
                     float eps2, CandBuf cand, hipStream_t st, int tile) {
    if (ngroups <= 0) return;
    tile = std::min(SOFT_TILE, std::max(64, tile & ~63));
    hipLaunchKernelGGL((k_soft_f32<IS, JS, SS, OUT, MODE>), dim3(xcd_grid(ngroups)), dim3(SOFT_BLOCK), 0,
                       st, d_groups, ngroups, is, js, jf, ss, out, eps2, cand, tile);
}
template <class IS, class JS, class SS, class OUT, int MODE>
void launch_soft_f64(const SoftGroup *d_groups, int ngroups, IS is, JS js, JS jf, SS ss, OUT out,
                     double eps2, CandBuf cand, hipStream_t st) {
    if (ngroups <= 0) return;
    hipLaunchKernelGGL((k_soft_f64<IS, JS, SS, OUT, MODE>), dim3(xcd_grid(ngroups)), dim3(SOFT_BLOCK), 0,
                       st, d_groups, ngroups, is, js, jf, ss, out, eps2, cand);
}

}  // namespace gplum
