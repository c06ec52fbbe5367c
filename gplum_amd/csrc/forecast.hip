// forecast.hip -- learned sizing of the EP-EP candidate buffer (see forecast.hpp).
#include <cmath>
#include <vector>

#include "forecast.hpp"

namespace gplum {

constexpr int FC_SCAN_BLOCK = 1024;

// Features of the particles of group g (one 64-thread block per group):
//  0  log density of the group's near list: (n_near + 1) / V, V = the group's
//     particle box grown by 1.1 max r_search on every side
//  1  log velocity dispersion, 2  log mass, 3  log Hill radius (m / 3 M_sun)^(1/3) |r|
__global__ __launch_bounds__(64) void k_fc_features(int ng, FcGroupView gv, const double *__restrict__ pos,
                                                    const double *__restrict__ mass,
                                                    const double *__restrict__ vdisp, double m_sun, FcNorm nrm,
                                                    int n, float *__restrict__ x) {
    const int g = blockIdx.x;
    if (g >= ng) return;
    const int node = gv.gnode[g];
    const double grow = 2.2 * gv.rs[node];
    double vol = 1.;
    for (int d = 0; d < 3; d++) vol *= (gv.bmax[3 * node + d] - gv.bmin[3 * node + d]) + grow;
    const float f0 = (float)log(((double)gv.ep_cnt[g] + 1.) / fmax(vol, 1e-300));
    const int i0 = gv.gstart[g], cnt = gv.gcount[g];
    for (int k = (int)threadIdx.x; k < cnt; k += 64) {
        const int i = i0 + k;
        const double m = mass[i];
        const double r = sqrt(pos[3 * i] * pos[3 * i] + pos[3 * i + 1] * pos[3 * i + 1] + pos[3 * i + 2] * pos[3 * i + 2]);
        const float f[MLP_NIN] = {f0, (float)log(fmax(vdisp[i], 1e-30)), (float)log(fmax(m, 1e-300)),
                                  (float)log(fmax(cbrt(m / (3. * m_sun)) * r, 1e-300))};
        for (int q = 0; q < MLP_NIN; q++) x[(size_t)q * n + i] = (f[q] - nrm.mu[q]) * nrm.inv_sd[q];
    }
}

__global__ void k_fc_normalize(int n, FcNorm nrm, float *__restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int q = 0; q < MLP_NIN; q++) x[(size_t)q * n + i] = (x[(size_t)q * n + i] - nrm.mu[q]) * nrm.inv_sd[q];
}

// capacity of group g: expected count S plus three Poisson standard deviations and 8;
// the rates are read in particle order over the group's range (uniform >= 0: every
// particle has that rate)
__global__ __launch_bounds__(64) void k_fc_capacity(int ng, const int32_t *__restrict__ gstart,
                                                    const int32_t *__restrict__ gcount,
                                                    const float *__restrict__ rate, float uniform,
                                                    int32_t *__restrict__ cap) {
    const int g = blockIdx.x;
    if (g >= ng) return;
    double s = 0.;
    const float *r = rate + gstart[g];
    for (int k = (int)threadIdx.x; k < gcount[g]; k += 64) s += (uniform >= 0.f) ? (double)uniform : (double)r[k];
    for (int d = 32; d > 0; d >>= 1) s += __shfl_xor(s, d);
    if (threadIdx.x == 0) cap[g] = (int32_t)fmin(ceil(s + 3. * sqrt(s) + 8.), 1e9);
}

// The one-block scans below: wave w of a pass covers FC_IPT * 64 consecutive groups, lane
// l the groups w * 64 * FC_IPT + k * 64 + l (k < FC_IPT), so that every load and store is
// coalesced; the exclusive offsets follow the group order (a wave scan per k with a
// running carry, then the waves' totals); a lane's FC_IPT loads are in flight together.
constexpr int FC_IPT = 8;

__device__ inline long long wave_incl_scan_ll(long long x) {
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const long long y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    return x;
}

// the waves' totals of a pass: *before = the waves before this one, returns the pass total
__device__ inline long long block_wave_totals(long long wave_total, long long *s_w, long long *before) {
    const int lane = __lane_id(), w = (int)threadIdx.x >> 6;
    constexpr int NW = FC_SCAN_BLOCK / 64;
    if (lane == 0) s_w[w] = wave_total;
    __syncthreads();
    long long b = 0, tot = 0;
    for (int j = 0; j < NW; j++) {
        b += j < w ? s_w[j] : 0;
        tot += s_w[j];
    }
    __syncthreads();
    *before = b;
    return tot;
}

__device__ inline long long wave_sum_ll(long long x) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d);
    return x;
}

// one block: exclusive offsets of the capacities (each at most FC_SORT_MAX, the keys a
// group sorts in LDS), clipped to the allocated slots.  Per pass: the waves' totals
// first, then each wave's scan in group order, written as it goes.
__global__ __launch_bounds__(FC_SCAN_BLOCK) void k_fc_offsets(int ng, int32_t *__restrict__ cap,
                                                              int32_t *__restrict__ off, long long slots,
                                                              int64_t *__restrict__ info) {
    __shared__ long long s_w[FC_SCAN_BLOCK / 64];
    const int lane = __lane_id(), w = (int)threadIdx.x >> 6;
    long long carry = 0;
    for (int base = 0; base < ng; base += FC_SCAN_BLOCK * FC_IPT) {
        const int g0 = base + w * 64 * FC_IPT + lane;
        int v[FC_IPT];
        long long sum = 0;
#pragma unroll
        for (int k = 0; k < FC_IPT; k++) {
            const int g = g0 + 64 * k;
            v[k] = g < ng ? min(cap[g], FC_SORT_MAX) : 0;
            sum += v[k];
        }
        long long before;
        const long long tot = block_wave_totals(wave_sum_ll(sum), s_w, &before);
        long long run = carry + before;
        for (int k = 0; k < FC_IPT; k++) {
            const long long x = wave_incl_scan_ll(v[k]);
            const long long o = run + x - v[k];
            run += __shfl(x, 63);
            const int g = g0 + 64 * k;
            if (g < ng) {
                off[g] = (int32_t)(o < slots ? o : slots);
                const long long room = slots - o;
                cap[g] = (int32_t)(room <= 0 ? 0 : ((long long)v[k] < room ? (long long)v[k] : room));
            }
        }
        carry += tot;
    }
    if (threadIdx.x == 0) info[0] = carry;
}

// one block: the final offsets of every group's candidates (exclusive scan of the exact
// counts: the groups' keys end up in group order, i.e. ascending i); the groups whose
// count exceeds their segment are listed, with the offsets of their candidates in the
// overflow area (in list order).  info[1..3] = total, overflowed groups, their total.
constexpr long long FC_PACK = 1ll << 40;
__global__ __launch_bounds__(FC_SCAN_BLOCK) void k_fc_summary(int ng, const int32_t *__restrict__ cnt,
                                                              const int32_t *__restrict__ cap,
                                                              int32_t *__restrict__ fin_off,
                                                              int32_t *__restrict__ ovf_list,
                                                              int32_t *__restrict__ ovf_pos,
                                                              int64_t *__restrict__ info) {
    __shared__ long long s_w[FC_SCAN_BLOCK / 64];
    const int lane = __lane_id(), w = (int)threadIdx.x >> 6;
    long long cs = 0, co = 0;
    for (int base = 0; base < ng; base += FC_SCAN_BLOCK * FC_IPT) {
        const int g0 = base + w * 64 * FC_IPT + lane;
        int c[FC_IPT];
        bool over[FC_IPT];
        long long sc = 0, so = 0;
#pragma unroll
        for (int k = 0; k < FC_IPT; k++) {
            const int g = g0 + 64 * k;
            c[k] = g < ng ? cnt[g] : 0;
            over[k] = g < ng && c[k] > cap[g];
            sc += c[k];
            // overflowed groups and their candidates in one sum (the group count above bit 40)
            so += over[k] ? FC_PACK + c[k] : 0;
        }
        long long bs, bo;
        const long long ts = block_wave_totals(wave_sum_ll(sc), s_w, &bs);
        const long long to = block_wave_totals(wave_sum_ll(so), s_w, &bo);
        long long rs = cs + bs, ro = co + bo;
        for (int k = 0; k < FC_IPT; k++) {
            const long long xs = wave_incl_scan_ll(c[k]);
            const long long es = rs + xs - c[k];
            rs += __shfl(xs, 63);
            // the overflow list's scan only where the wave has an overflowed group (rare)
            long long eo = ro;
            if (__ballot(over[k])) {
                const long long vo = over[k] ? FC_PACK + c[k] : 0;
                const long long xo = wave_incl_scan_ll(vo);
                eo = ro + xo - vo;
                ro += __shfl(xo, 63);
            }
            const int g = g0 + 64 * k;
            if (g < ng) {
                fin_off[g] = (int32_t)es;
                if (over[k]) {
                    ovf_list[eo / FC_PACK] = g;
                    ovf_pos[eo / FC_PACK] = (int32_t)(eo % FC_PACK);
                }
            }
        }
        cs += ts; co += to;
    }
    if (threadIdx.x == 0) { info[1] = cs; info[2] = co / FC_PACK; info[3] = co % FC_PACK; }
}

// The candidates of a group that fits its segment, as sort keys i 2^kb + j in ascending
// order (bitonic sort in LDS), at the group's final offset.
__global__ __launch_bounds__(FC_SORT_BLOCK) void k_fc_sort_keys(int ng, const int32_t *__restrict__ cnt,
                                                                const int32_t *__restrict__ cap,
                                                                const int32_t *__restrict__ seg_off,
                                                                const int32_t *__restrict__ fin_off,
                                                                const int2 *__restrict__ seg, int kb,
                                                                uint64_t *__restrict__ keys, long long keys_cap) {
    __shared__ uint64_t s_k[FC_SORT_MAX];
    const int g = blockIdx.x;
    if (g >= ng) return;
    const int c = cnt[g];
    // (a group whose keys would pass the end of the key array is left out: the caller
    // launches the sort again on a larger array)
    if (c == 0 || c > cap[g] || (long long)fin_off[g] + c > keys_cap) return;
    int P = 1;
    while (P < c) P <<= 1;
    const int2 *src = seg + seg_off[g];
    for (int k = (int)threadIdx.x; k < P; k += FC_SORT_BLOCK) {
        uint64_t v = ~0ull;
        if (k < c) {
            const int2 q = src[k];
            v = ((uint64_t)(uint32_t)q.x << kb) | (uint64_t)(uint32_t)q.y;
        }
        s_k[k] = v;
    }
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = (int)threadIdx.x; t < P / 2; t += FC_SORT_BLOCK) {
                const int lo = 2 * stride * (t / stride) + (t % stride), hi = lo + stride;
                const bool up = (lo & size) == 0;
                const uint64_t a = s_k[lo], b = s_k[hi];
                if ((a > b) == up) { s_k[lo] = b; s_k[hi] = a; }
            }
            __syncthreads();
        }
    uint64_t *dst = keys + fin_off[g];
    for (int k = (int)threadIdx.x; k < c; k += FC_SORT_BLOCK) dst[k] = s_k[k];
}

// The overflowed groups' candidates, sorted together (ascending i, so group after group
// in list order), copied to each group's final offset.
__global__ void k_fc_scatter_ovf(int n_ovf, const int32_t *__restrict__ ovf_list,
                                 const int32_t *__restrict__ ovf_pos, const int32_t *__restrict__ cnt,
                                 const int32_t *__restrict__ fin_off, const uint64_t *__restrict__ sorted,
                                 uint64_t *__restrict__ keys) {
    const int k = blockIdx.x;
    if (k >= n_ovf) return;
    const int g = ovf_list[k];
    const int c = cnt[g];
    for (int t = (int)threadIdx.x; t < c; t += blockDim.x) keys[fin_off[g] + t] = sorted[ovf_pos[k] + t];
}

__global__ void k_fc_gather_groups(int m, const int32_t *__restrict__ list, const SoftGroup *__restrict__ groups,
                                   SoftGroup *__restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) out[k] = groups[list[k]];
}

__global__ void k_fc_sample(int m, int n, long long stride, long long phase, const float *__restrict__ x,
                            const int32_t *__restrict__ cnt, float *__restrict__ xs, float *__restrict__ ys) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m) return;
    const int i = (int)((phase + (long long)s * stride) % n);
    for (int q = 0; q < MLP_NIN; q++) xs[(size_t)q * m + s] = x[(size_t)q * n + i];
    ys[s] = (float)cnt[i];
}

// ------------------------------------------------------------------ host

int fc_features(CandSizer &s, const FcGroupView &gv, int ng, const double *pos, const double *mass,
                const double *vdisp, double m_sun, int n, hipStream_t st) {
    GPLUM_ENSURE(s.x, (size_t)MLP_NIN * n);
    FcNorm nrm = s.normalized ? s.norm : FcNorm{{0.f, 0.f, 0.f, 0.f}, {1.f, 1.f, 1.f, 1.f}};
    if (ng > 0)
        hipLaunchKernelGGL(k_fc_features, dim3(ng), dim3(64), 0, st, ng, gv, pos, mass, vdisp, m_sun, nrm, n,
                           s.x.p);
    GPLUM_HIP(hipGetLastError());
    return GPLUM_OK;
}

int fc_plan(CandSizer &s, const FcGroupView &gv, int ng, int n, float *rate, bool fresh, CandBuf *cb,
            hipStream_t st) {
    GPLUM_ENSURE(s.seg_off, ng + 1);
    GPLUM_ENSURE(s.seg_cap, ng + 1);
    GPLUM_ENSURE(s.seg_cnt, ng + 1);
    GPLUM_ENSURE(s.info, 4);
    if (s.seg_capacity > (size_t)INT32_MAX) s.seg_capacity = INT32_MAX;
    GPLUM_ENSURE(s.pairs_seg, s.seg_capacity + 1);
    if (!s.uniform && (fresh || !s.lam_valid)) {
        int rc = s.net.predict(s.x.p, n, rate, st);
        if (rc) return rc;
        s.lam_valid = true;
    }
    if (ng > 0) {
        hipLaunchKernelGGL(k_fc_capacity, dim3(ng), dim3(64), 0, st, ng, gv.gstart, gv.gcount, rate,
                           s.uniform ? s.uniform_rate : -1.f, s.seg_cap.p);
        hipLaunchKernelGGL(k_fc_offsets, dim3(1), dim3(FC_SCAN_BLOCK), 0, st, ng, s.seg_cap.p, s.seg_off.p,
                           (long long)s.seg_capacity, s.info.p);
    }
    GPLUM_HIP(hipGetLastError());
    CandBuf c{s.pairs_seg.p, nullptr, 0u};
    c.seg_off = s.seg_off.p;
    c.seg_cap = s.seg_cap.p;
    c.seg_cnt = s.seg_cnt.p;
    *cb = c;
    return GPLUM_OK;
}

int fc_summary(CandSizer &s, int ng, hipStream_t st) {
    GPLUM_ENSURE(s.fin_off, ng + 1);
    GPLUM_ENSURE(s.ovf_list, ng + 1);
    GPLUM_ENSURE(s.ovf_pos, ng + 1);
    if (ng > 0)
        hipLaunchKernelGGL(k_fc_summary, dim3(1), dim3(FC_SCAN_BLOCK), 0, st, ng, s.seg_cnt.p, s.seg_cap.p,
                           s.fin_off.p, s.ovf_list.p, s.ovf_pos.p, s.info.p);
    GPLUM_HIP(hipGetLastError());
    return GPLUM_OK;
}

int fc_read_info(CandSizer &s, int64_t out[4], hipStream_t st) {
    GPLUM_HIP(hipMemcpyAsync(out, s.info.p, sizeof(int64_t) * 4, hipMemcpyDeviceToHost, st));
    GPLUM_HIP(hipStreamSynchronize(st));
    return GPLUM_OK;
}

int fc_sort_keys(CandSizer &s, int ng, int kb, uint64_t *keys, long long keys_cap, hipStream_t st) {
    if (ng > 0)
        hipLaunchKernelGGL(k_fc_sort_keys, dim3(ng), dim3(FC_SORT_BLOCK), 0, st, ng, s.seg_cnt.p, s.seg_cap.p,
                           s.seg_off.p, s.fin_off.p, s.pairs_seg.p, kb, keys, keys_cap);
    GPLUM_HIP(hipGetLastError());
    return GPLUM_OK;
}

int fc_scatter_overflow(CandSizer &s, int n_ovf, const uint64_t *sorted, uint64_t *keys, hipStream_t st) {
    if (n_ovf > 0)
        hipLaunchKernelGGL(k_fc_scatter_ovf, dim3(n_ovf), dim3(256), 0, st, n_ovf, s.ovf_list.p, s.ovf_pos.p,
                           s.seg_cnt.p, s.fin_off.p, sorted, keys);
    GPLUM_HIP(hipGetLastError());
    return GPLUM_OK;
}

int fc_overflow_groups(CandSizer &s, const SoftGroup *groups, int n_ovf, hipStream_t st) {
    GPLUM_ENSURE(s.ovf_groups, n_ovf + 1);
    if (n_ovf > 0)
        hipLaunchKernelGGL(k_fc_gather_groups, dim3(ceil_div(n_ovf, 256)), dim3(256), 0, st, n_ovf, s.ovf_list.p,
                           groups, s.ovf_groups.p);
    GPLUM_HIP(hipGetLastError());
    return GPLUM_OK;
}

int fc_train(CandSizer &s, const int32_t *count_fp32, int n, hipStream_t st, bool want_loss) {
    if (n <= 0) return GPLUM_OK;
    int rc = s.net.init_params(st, 12345u);
    if (rc) return rc;
    if (!s.normalized) {
        // standardise the raw features of this first step (host, once)
        std::vector<float> h((size_t)MLP_NIN * n);
        GPLUM_HIP(hipMemcpyAsync(h.data(), s.x.p, sizeof(float) * h.size(), hipMemcpyDeviceToHost, st));
        GPLUM_HIP(hipStreamSynchronize(st));
        for (int q = 0; q < MLP_NIN; q++) {
            double a = 0., b = 0.;
            for (int i = 0; i < n; i++) a += h[(size_t)q * n + i];
            a /= n;
            for (int i = 0; i < n; i++) {
                const double d = h[(size_t)q * n + i] - a;
                b += d * d;
            }
            const double sd = std::sqrt(b / n);
            s.norm.mu[q] = (float)a;
            s.norm.inv_sd[q] = (float)(sd > 1e-6 ? 1. / sd : 1.);
        }
        hipLaunchKernelGGL(k_fc_normalize, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, s.norm, s.x.p);
        s.normalized = true;
    }
    const int m = n < s.batch ? n : s.batch;
    const long long stride = n / m;
    GPLUM_ENSURE(s.xs, (size_t)MLP_NIN * m);
    GPLUM_ENSURE(s.ys, m);
    for (int k = 0; k < s.train_steps; k++) {
        const long long phase = (long long)((s.phase++ * 2654435761u) % (unsigned)(stride > 0 ? stride : 1));
        hipLaunchKernelGGL(k_fc_sample, dim3(ceil_div(m, 256)), dim3(256), 0, st, m, n, stride, phase, s.x.p,
                           count_fp32, s.xs.p, s.ys.p);
        double *lp = (want_loss && k == s.train_steps - 1) ? &s.last_loss : nullptr;
        if ((rc = s.net.loss_grad(s.xs.p, s.ys.p, m, lp, st))) return rc;
        if ((rc = s.net.adam_step(st))) return rc;
    }
    s.last_batch = m;
    s.trained_rounds++;
    return GPLUM_OK;
}

int fc_read_loss(CandSizer &s, hipStream_t st) {
    if (s.last_batch <= 0) return GPLUM_OK;
    return s.net.read_loss(s.last_batch, &s.last_loss, st);
}

}  // namespace gplum
