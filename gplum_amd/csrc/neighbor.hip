// neighbor.hip -- FP64 neighbour decision + changeover correction, neighbour CSR,
// connected components and cluster segmentation.
//
// Candidates are the (i, j) pairs the FP32 EP-EP kernel flagged with
// |r|^2+eps^2 < 1.0201 max(r_search)^2.  They are sorted by (i, j); row i then holds
// exactly neighbor.number entries, the same candidate set correctForceLong examines
// (src/gravity_soft.h:283-312: id_min/id_max for <= 2, the FDPS query otherwise;
// every pair closer than max(r_search) is a candidate either way).  One thread per
// particle applies correctForceBetween2Particles[Initial] (:75-242) to its row in
// ascending j, which fixes the FP64 summation order.
#include <cfloat>
#include <hipcub/hipcub.hpp>

#include "neighbor.hpp"

namespace gplum {

struct CorrAcc {
    double phi, phi_d, acc[3], acc_d[3], jerk[3], acc0;
    int number;
    int64_t id_cluster;
};

struct PView {
    const double *pos, *vel, *acc_d, *mass, *r_out, *r_out_inv, *r_search;
    const int64_t *id;
};

// correctForceBetween2Particles[Initial], same operation order as the oracle.
__device__ inline bool correct_pair(const NbParams &P, const PView &V, int i, int j, bool initial,
                                    CorrAcc &a) {
    const double massj = V.mass[j];
    const double r_out = smax(V.r_out[i], V.r_out[j]);
    const double r_out_inv = smin(V.r_out_inv[i], 1. / V.r_out[j]);
    const double r_search = smax(V.r_search[i], V.r_search[j]);
    bool accepted = false;
    const double drx = V.pos[3 * j] - V.pos[3 * i];
    const double dry = V.pos[3 * j + 1] - V.pos[3 * i + 1];
    const double drz = V.pos[3 * j + 2] - V.pos[3 * i + 2];
    double dr2 = drx * drx + dry * dry + drz * drz;
    dr2 += P.eps2;
    const double rij = sqrt(dr2);
    const double dvx = V.vel[3 * j] - V.vel[3 * i];
    const double dvy = V.vel[3 * j + 1] - V.vel[3 * i + 1];
    const double dvz = V.vel[3 * j + 2] - V.vel[3 * i + 2];
    const double dax = V.acc_d[3 * j] - V.acc_d[3 * i];
    const double day = V.acc_d[3 * j + 1] - V.acc_d[3 * i + 1];
    const double daz = V.acc_d[3 * j + 2] - V.acc_d[3 * i + 2];
    const double drdv = drx * dvx + dry * dvy + drz * dvz;
    const double dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
    const double da2 = dax * dax + day * day + daz * daz;
    // the reference's t_min (units of length) with its std::min/std::max order: NaN when
    // dv2 == 0, and dr2_min then falls back to dr2
    const double t_min = smin(smax(-drdv / sqrt(dv2), 0.), P.dt_tree);
    const double dr2_min = smin(dr2, dr2 + 2. * drdv * t_min + dv2 * t_min * t_min);
    const double r_crit = P.R_search2 * r_out;
    const double v_crit_a = P.R_search3 * 0.5 * P.dt_tree;
    bool cond = (dr2_min < r_crit * r_crit) || (dv2 < v_crit_a * v_crit_a * da2);
    if (initial) cond = cond || (da2 == 0.);
    if (cond && rij < r_search) {
        a.number += 1;
        if (V.id[j] < a.id_cluster) a.id_cluster = V.id[j];
        a.acc0 += r_out * r_out / massj;
        accepted = true;
    }
    if (rij < r_out) {
        const double rinv = 1. / rij;
        const double r2inv = rinv * rinv;
        const double r3inv = rinv * r2inv;
        const double W = cutoff_W(P.cut, rij, r_out_inv);
        const double K = cutoff_K(P.cut, rij, r_out_inv);
        const double r_min = smin(rinv, r_out_inv);
        a.phi -= massj * (rinv * W - r_min);
        const double cf = massj * (r3inv * K - r_min * r_min * r_min);
        a.acc[0] += cf * drx;
        a.acc[1] += cf * dry;
        a.acc[2] += cf * drz;
        if (initial) {
            const double alpha = drdv * r2inv;
            const double dKdt = cutoff_dKdt(P.cut, rij, r_out_inv, alpha);
            const double alpha_c = alpha * (1. - K);
            a.phi_d -= massj * rinv * (1. - W);
            const double cd = massj * r3inv * (1. - K);
            a.acc_d[0] += cd * drx;
            a.acc_d[1] += cd * dry;
            a.acc_d[2] += cd * drz;
            const double mr3 = massj * r3inv;
            const double c2 = 3. * alpha_c + dKdt;
            a.jerk[0] += mr3 * ((1. - K) * dvx - c2 * drx);
            a.jerk[1] += mr3 * ((1. - K) * dvy - c2 * dry);
            a.jerk[2] += mr3 * ((1. - K) * dvz - c2 * drz);
        }
    }
    return accepted;
}

__device__ inline double calc_dt2(double eta, double alpha2, double acc0, const double a[3],
                                  const double j[3]) {
    const double Acc2 = dot3(a[0], a[1], a[2], a[0], a[1], a[2]) + alpha2 * acc0 * acc0;
    const double Jerk2 = dot3(j[0], j[1], j[2], j[0], j[1], j[2]);
    return (Jerk2 > 0.) ? eta * sqrt(Acc2 / Jerk2) : DBL_MAX;
}

// candidate (i, j) -> sort key i 2^kb + j (kb bits hold any particle index)
__global__ void k_cand_keys(const int2 *__restrict__ cand, int nc, int kb, uint64_t *__restrict__ key) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int2 q = cand[c];
    key[c] = ((uint64_t)(uint32_t)q.x << kb) | (uint64_t)(uint32_t)q.y;
}

__global__ void k_finalize(NbParams P, PView V, int n, int ncand, int initial, uint64_t jmask,
                           const int32_t *__restrict__ row_off,
                           const int32_t *__restrict__ number_fp32, const uint64_t *__restrict__ keys,
                           unsigned char *__restrict__ accepted, double *__restrict__ acc,
                           double *__restrict__ phi, double *__restrict__ acc0, int32_t *__restrict__ number,
                           int64_t *__restrict__ id_cluster, double *__restrict__ acc_d_new,
                           double *__restrict__ jerk_d, double *__restrict__ phi_d,
                           double *__restrict__ acc_s, double *__restrict__ jerk_s,
                           double *__restrict__ phi_s, double *__restrict__ dt,
                           const double *__restrict__ time) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    CorrAcc a;
    a.phi = a.phi_d = a.acc0 = 0.;
    for (int k = 0; k < 3; k++) a.acc[k] = a.acc_d[k] = a.jerk[k] = 0.;
    a.number = 0;
    a.id_cluster = V.id[i];
    a.phi += V.mass[i] * V.r_out_inv[i];     // self term (gravity_soft.h:280, 292)
    // rows end at ncand even if the FP32 counts and the candidates found disagree (an
    // error the host reports after the stage)
    const int b = row_off[i], e = min(b + number_fp32[i], ncand);
    for (int c = b; c < e; c++) {
        const int j = (int)(keys[c] & jmask);
        bool acc_flag = false;
        if (V.id[j] != V.id[i]) acc_flag = correct_pair(P, V, i, j, initial != 0, a);
        accepted[c] = acc_flag ? 1 : 0;
    }
    acc[3 * i] += a.acc[0];
    acc[3 * i + 1] += a.acc[1];
    acc[3 * i + 2] += a.acc[2];
    phi[i] += a.phi;
    acc0[i] = (a.acc0 > 0.) ? (double)a.number / a.acc0 : 0.;
    number[i] = a.number;
    id_cluster[i] = a.id_cluster;
    if (initial) {
        // calcStarGravity (gravity_hard.h:5-39)
        const double x[3] = {V.pos[3 * i], V.pos[3 * i + 1], V.pos[3 * i + 2]};
        const double v[3] = {V.vel[3 * i], V.vel[3 * i + 1], V.vel[3 * i + 2]};
        const double dr[3] = {-x[0], -x[1], -x[2]};
        const double dv[3] = {-v[0], -v[1], -v[2]};
        const double r2inv = 1. / (dot3(dr[0], dr[1], dr[2], dr[0], dr[1], dr[2]) + P.eps2_sun);
        const double rinv = sqrt(r2inv);
        const double r3inv = rinv * r2inv;
        const double mj = P.m_sun * r3inv;
        const double alpha = dot3(dr[0], dr[1], dr[2], dv[0], dv[1], dv[2]) * r2inv;
        double as[3], js[3];
        for (int k = 0; k < 3; k++) {
            as[k] = mj * dr[k];
            js[k] = mj * (dv[k] - 3. * alpha * dr[k]);
            acc_s[3 * i + k] = as[k];
            jerk_s[3 * i + k] = js[k];
            acc_d_new[3 * i + k] = a.acc_d[k];
            jerk_d[3 * i + k] = a.jerk[k];
        }
        phi_s[i] = -P.m_sun * rinv;
        phi_d[i] = a.phi_d;
        // calcDeltatInitial runs before acc_d is written back (gravity_soft.h:502-516)
        const double ad_old[3] = {V.acc_d[3 * i], V.acc_d[3 * i + 1], V.acc_d[3 * i + 2]};
        const double a0 = acc0[i];
        double dt_next = 0.5 * P.dt_tree;
        const double dt_1 = smin(calc_dt2(P.eta_0, P.alpha2, a0, ad_old, a.jerk),
                                 calc_dt2(P.eta_sun0, P.alpha2, 0., as, js));
        const double tm = time[i];
        double rem = fmod_block(tm, dt_next);
        for (int it = 0; rem != 0.0 && it < HALVE_CAP; it++) { dt_next *= 0.5; rem = fmod_block(tm, dt_next); }
        const double dt0 = dt[i];
        if (dt0 > 0.)
            for (int it = 0; 2. * dt0 < dt_next && it < HALVE_CAP; it++) dt_next *= 0.5;
        for (int it = 0; dt_1 < dt_next && it < HALVE_CAP; it++) dt_next *= 0.5;
        if (dt_next < 2. * P.dt_min || rem != 0.0) dt_next = P.dt_min;
        dt[i] = dt_next;
    }
}

__global__ void k_compact_rows(int n, int ncand, uint64_t jmask, const int32_t *__restrict__ row_off,
                               const int32_t *__restrict__ number_fp32, const uint64_t *__restrict__ keys,
                               const unsigned char *__restrict__ accepted, const int32_t *__restrict__ nbr_off,
                               int32_t *__restrict__ nbr_idx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int w = nbr_off[i];
    const int b = row_off[i], e = min(b + number_fp32[i], ncand);
    for (int c = b; c < e; c++)
        if (accepted[c]) nbr_idx[w++] = (int)(keys[c] & jmask);
}

// ---------------------------------------------------------- union-find

__device__ inline int uf_find(const int32_t *parent, int x) {
    int p = __atomic_load_n(&parent[x], __ATOMIC_RELAXED);
    while (p != x) {
        x = p;
        p = __atomic_load_n(&parent[x], __ATOMIC_RELAXED);
    }
    return x;
}

__global__ void k_uf_init(int n, int32_t *parent) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) parent[i] = i;
}

// Link roots so that the root of every component is its minimum-id member; the
// fixed point of createNeighborCluster (neighbor.h:799-839) is that id.
__global__ void k_uf_hook(int n, const int32_t *__restrict__ nbr_off, const int32_t *__restrict__ nbr_idx,
                          const int64_t *__restrict__ id, int32_t *parent) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int e = nbr_off[i]; e < nbr_off[i + 1]; e++) {
        const int j = nbr_idx[e];
        if (!(i < j)) continue;
        while (true) {
            const int ri = uf_find(parent, i), rj = uf_find(parent, j);
            if (ri == rj) break;
            const bool ri_hi = id[ri] > id[rj];
            const int hi = ri_hi ? ri : rj, lo = ri_hi ? rj : ri;
            if (atomicCAS(&parent[hi], hi, lo) == hi) break;
        }
    }
}

__global__ void k_uf_label(int n, const int32_t *__restrict__ number, const int64_t *__restrict__ id,
                           int32_t *parent, int64_t *__restrict__ id_cluster, uint64_t *__restrict__ ckey,
                           int32_t *__restrict__ cflag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = uf_find(parent, i);
    const int64_t lab = (number[i] > 0) ? id[r] : id[i];
    id_cluster[i] = lab;
    ckey[i] = (uint64_t)lab;
    cflag[i] = number[i] > 0 ? 1 : 0;
}

__global__ void k_select(int n, const int32_t *__restrict__ flag, const int32_t *__restrict__ pos,
                         const uint64_t *__restrict__ key, uint64_t *__restrict__ key_out,
                         int32_t *__restrict__ idx_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    key_out[pos[i]] = key[i];
    idx_out[pos[i]] = i;
}

__global__ void k_segment_flags(int m, const uint64_t *__restrict__ key, int32_t *__restrict__ flag) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    flag[k] = (k == 0 || key[k] != key[k - 1]) ? 1 : 0;
}

// segment starts -> cluster offsets (off holds m + 1 entries, enough for any count);
// info[0] = the number of clusters
__global__ void k_segment_offsets(int m, const int32_t *__restrict__ flag, const int32_t *__restrict__ scan,
                                  int32_t *__restrict__ off, int32_t *__restrict__ info) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > m) return;
    if (k == m) {
        const int ncl = scan[m - 1] + flag[m - 1];
        off[ncl] = m;
        info[0] = ncl;
        return;
    }
    if (flag[k]) off[scan[k]] = k;
}

// info[1] = the largest cluster (info[1] cleared before)
__global__ void k_segment_max(int m, const int32_t *__restrict__ flag, const int32_t *__restrict__ scan,
                              const int32_t *__restrict__ off, int32_t *__restrict__ info) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m || !flag[k]) return;
    const int c = scan[k];
    atomicMax(&info[1], off[c + 1] - off[c]);
}

// Members of each cluster in ascending id (the reference's particle-array order on one
// rank): rank of member a inside its cluster, then scatter.  Clusters are small, so
// the O(n_c^2) rank count costs less than a segmented sort.
__global__ void k_order_by_id(int m, const int32_t *__restrict__ flag, const int32_t *__restrict__ scan,
                              const int32_t *__restrict__ off, const int32_t *__restrict__ index,
                              const int64_t *__restrict__ id, int32_t *__restrict__ index_out) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= m) return;
    const int c = scan[a] + flag[a] - 1;
    const int b0 = off[c], b1 = off[c + 1];
    const int ia = index[a];
    const int64_t ida = id[ia];
    int r = 0;
    for (int b = b0; b < b1; b++) {
        const int64_t idb = id[index[b]];
        r += (idb < ida || (idb == ida && b < a)) ? 1 : 0;
    }
    index_out[b0 + r] = ia;
}

// Processing order of the clusters: the costliest first, the cost estimated as block
// steps (dt_tree over the smallest current step of the members) times (4 + size), so that
// the cluster with the longest chain of block steps starts at once rather than after the
// bulk of the small ones.
__global__ void k_cluster_cost(int ncl, const int32_t *__restrict__ off, const int32_t *__restrict__ index,
                               const double *__restrict__ dt, double dt_tree, int32_t *__restrict__ key,
                               int32_t *__restrict__ idx) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncl) return;
    const int b = off[c], e = off[c + 1], sz = e - b;
    double dmin = dt_tree;
    for (int a = b; a < e; a++) {
        const double d = dt[index[a]];
        if (d > 0. && d < dmin) dmin = d;
    }
    const double cost = (dt_tree / dmin) * (4.0 + sz);
    key[c] = -(int32_t)fmin(2.0e9, 64.0 * log2(cost));   // ascending sort: costliest first
    idx[c] = c;
}

// ------------------------------------------------------------------ host

template <class T>
static int nb_scan(Scratch &s, const T *in, T *out, int n, hipStream_t st) {
    size_t bytes = 0;
    GPLUM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, st));
    GPLUM_ENSURE(s.tmp, bytes + 16);
    GPLUM_HIP(hipcub::DeviceScan::ExclusiveSum(s.tmp.p, bytes, in, out, n, st));
    return GPLUM_OK;
}

template <class T>
static int last_sum(const T *in, const T *scan, int n, T *total, hipStream_t st);

// info[0] = 1 if the FP32 counts do not add up to the candidates found; nbr_off[n] and
// info[1] = the accepted total (the last offset plus the last count)
__global__ void k_finalize_totals(int n, int ncand, const int32_t *__restrict__ number_fp32,
                                  const int32_t *__restrict__ row_off, const int32_t *__restrict__ number,
                                  int32_t *__restrict__ nbr_off, int32_t *__restrict__ info) {
    if (threadIdx.x != 0) return;
    const int total = n > 0 ? row_off[n - 1] + number_fp32[n - 1] : 0;
    info[0] = (total != ncand) ? 1 : 0;
    const int nn = n > 0 ? nbr_off[n - 1] + number[n - 1] : 0;
    nbr_off[n] = nn;
    info[1] = nn;
}

template <class T>
static int last_sum(const T *in, const T *scan, int n, T *total, hipStream_t st) {
    T a = 0, b = 0;
    if (n <= 0) { *total = 0; return GPLUM_OK; }
    GPLUM_HIP(hipMemcpyAsync(&a, in + n - 1, sizeof(T), hipMemcpyDeviceToHost, st));
    GPLUM_HIP(hipMemcpyAsync(&b, scan + n - 1, sizeof(T), hipMemcpyDeviceToHost, st));
    GPLUM_HIP(hipStreamSynchronize(st));
    *total = a + b;
    return GPLUM_OK;
}

NbParams make_nb_params(const gplum_params &p) {
    NbParams q;
    q.eps2 = p.eps2; q.eps2_sun = p.eps2_sun; q.m_sun = p.m_sun; q.dt_tree = p.dt_tree;
    q.dt_min = p.dt_min; q.R_search2 = p.R_search2; q.R_search3 = p.R_search3;
    q.eta_0 = p.eta_0; q.eta_sun0 = p.eta_sun0; q.alpha2 = p.alpha2; q.cut = cutoff_of(p);
    return q;
}

int cand_key_bits(int n) {
    int kb = 1;
    while (kb < 31 && (1ll << kb) < (long long)n) kb++;
    return kb;
}

int cand_sort(const int2 *d_cand, int ncand, int kb, DArray<uint64_t> &keys_in, uint64_t *keys_out, Scratch &s,
              hipStream_t st) {
    if (ncand <= 0) return GPLUM_OK;
    GPLUM_ENSURE(keys_in, ncand + 1);
    hipLaunchKernelGGL(k_cand_keys, dim3(ceil_div(ncand, 256)), dim3(256), 0, st, d_cand, ncand, kb, keys_in.p);
    size_t bytes = 0;
    GPLUM_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, keys_in.p, keys_out, ncand, 0, 2 * kb, st));
    GPLUM_ENSURE(s.tmp, bytes + 16);
    GPLUM_HIP(hipcub::DeviceRadixSort::SortKeys(s.tmp.p, bytes, keys_in.p, keys_out, ncand, 0, 2 * kb, st));
    return GPLUM_OK;
}

int neighbor_finalize(const gplum_params &p, ParticleSoA &ps, int n, const int2 *d_cand, int ncand,
                      bool initial, NbState &nb, Scratch &s, hipStream_t st, bool presorted) {
    GPLUM_ENSURE(nb.keys, ncand + 1);
    GPLUM_ENSURE(nb.accepted, ncand + 1);
    GPLUM_ENSURE(nb.row_off, n + 1);
    GPLUM_ENSURE(nb.nbr_off, n + 1);
    GPLUM_ENSURE(nb.acc_d_new, 3 * (size_t)n);
    const int kb = cand_key_bits(n);
    const uint64_t jmask = (1ull << kb) - 1ull;
    if (!presorted) {
        int rc0 = cand_sort(d_cand, ncand, kb, nb.keys_in, nb.keys.p, s, st);
        if (rc0) return rc0;
    }
    int rc = nb_scan(s, ps.number.p, nb.row_off.p, n, st);
    if (rc) return rc;
    // number (FP32 count) is overwritten by the FP64 count: keep a copy
    GPLUM_ENSURE(nb.number_fp32, n);
    GPLUM_HIP(hipMemcpyAsync(nb.number_fp32.p, ps.number.p, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st));
    NbParams P = make_nb_params(p);
    PView V{ps.pos.p, ps.vel.p, ps.acc_d.p, ps.mass.p, ps.r_out.p, ps.r_out_inv.p, ps.r_search.p, ps.id.p};
    hipLaunchKernelGGL(k_finalize, dim3(ceil_div(n, 128)), dim3(128), 0, st, P, V, n, ncand, initial ? 1 : 0,
                       jmask, nb.row_off.p, nb.number_fp32.p, nb.keys.p, nb.accepted.p, ps.acc.p, ps.phi.p,
                       ps.acc0.p, ps.number.p, ps.id_cluster.p, nb.acc_d_new.p, ps.jerk_d.p, ps.phi_d.p,
                       ps.acc_s.p, ps.jerk_s.p, ps.phi_s.p, ps.dt.p, ps.time.p);
    if (initial)
        GPLUM_HIP(hipMemcpyAsync(ps.acc_d.p, nb.acc_d_new.p, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, st));
    // accepted-neighbour CSR (accepted <= candidates: the index array is sized by the
    // candidates, so nothing waits for the accepted total before the compaction)
    if ((rc = nb_scan(s, ps.number.p, nb.nbr_off.p, n, st))) return rc;
    GPLUM_ENSURE(nb.info, 2);
    hipLaunchKernelGGL(k_finalize_totals, dim3(1), dim3(64), 0, st, n, ncand, nb.number_fp32.p, nb.row_off.p,
                       ps.number.p, nb.nbr_off.p, nb.info.p);
    GPLUM_ENSURE(nb.nbr_idx, (size_t)ncand + 1);
    hipLaunchKernelGGL(k_compact_rows, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, ncand, jmask, nb.row_off.p,
                       nb.number_fp32.p, nb.keys.p, nb.accepted.p, nb.nbr_off.p, nb.nbr_idx.p);
    GPLUM_HIP(hipGetLastError());
    // the check and the accepted total are read back without a wait: nb_resolve takes
    // them at the caller's next synchronisation
    GPLUM_HIP(nb.info_host.ensure(2 * sizeof(int32_t)));
    GPLUM_HIP(hipMemcpyAsync(nb.info_host.p, nb.info.p, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    nb.info_pending = true;
    nb.n = n;
    return GPLUM_OK;
}

int nb_resolve(NbState &nb) {
    if (!nb.info_pending) return GPLUM_OK;
    nb.info_pending = false;
    const int32_t *info = nb.info_host.as<int32_t>();
    if (info[0]) return fail(GPLUM_E_STATE, "candidate count mismatch (FP32 counts vs candidates found)");
    nb.n_nbr = info[1];
    return GPLUM_OK;
}

int make_clusters(ParticleSoA &ps, int n, NbState &nb, ClusterList &cl, Scratch &s, double dt_tree,
                  hipStream_t st) {
    GPLUM_ENSURE(cl.parent, n);
    GPLUM_ENSURE(cl.ckey, n);
    GPLUM_ENSURE(cl.cflag, n);
    GPLUM_ENSURE(cl.cscan, n);
    hipLaunchKernelGGL(k_uf_init, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, cl.parent.p);
    hipLaunchKernelGGL(k_uf_hook, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, nb.nbr_off.p, nb.nbr_idx.p,
                           ps.id.p, cl.parent.p);
    hipLaunchKernelGGL(k_uf_label, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, ps.number.p, ps.id.p,
                       cl.parent.p, ps.id_cluster.p, cl.ckey.p, cl.cflag.p);
    int rc = nb_scan(s, cl.cflag.p, cl.cscan.p, n, st);
    if (rc) return rc;
    int m = 0;
    if ((rc = last_sum(cl.cflag.p, cl.cscan.p, n, &m, st))) return rc;
    if ((rc = nb_resolve(nb))) return rc;       // the stream has passed the finalize
    cl.m = m;
    cl.ncl = 0;
    if (m == 0) return GPLUM_OK;
    GPLUM_ENSURE(cl.key_sel, m);
    GPLUM_ENSURE(cl.idx_sel, m);
    GPLUM_ENSURE(cl.key_sorted, m);
    GPLUM_ENSURE(cl.index, m);
    hipLaunchKernelGGL(k_select, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, cl.cflag.p, cl.cscan.p,
                       cl.ckey.p, cl.key_sel.p, cl.idx_sel.p);
    size_t bytes = 0;
    GPLUM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, cl.key_sel.p, cl.key_sorted.p, cl.idx_sel.p,
                                                  cl.index.p, m, 0, 64, st));
    GPLUM_ENSURE(s.tmp, bytes + 16);
    GPLUM_HIP(hipcub::DeviceRadixSort::SortPairs(s.tmp.p, bytes, cl.key_sel.p, cl.key_sorted.p, cl.idx_sel.p,
                                                  cl.index.p, m, 0, 64, st));
    GPLUM_ENSURE(cl.sflag, m);
    GPLUM_ENSURE(cl.sscan, m);
    hipLaunchKernelGGL(k_segment_flags, dim3(ceil_div(m, 256)), dim3(256), 0, st, m, cl.key_sorted.p, cl.sflag.p);
    if ((rc = nb_scan(s, cl.sflag.p, cl.sscan.p, m, st))) return rc;
    // offsets, cluster count and largest cluster, read back together
    GPLUM_ENSURE(cl.off, (size_t)m + 1);
    GPLUM_ENSURE(cl.max_dev, 2);
    GPLUM_HIP(hipMemsetAsync(cl.max_dev.p, 0, 2 * sizeof(int32_t), st));
    hipLaunchKernelGGL(k_segment_offsets, dim3(ceil_div(m + 1, 256)), dim3(256), 0, st, m, cl.sflag.p,
                       cl.sscan.p, cl.off.p, cl.max_dev.p);
    hipLaunchKernelGGL(k_segment_max, dim3(ceil_div(m, 256)), dim3(256), 0, st, m, cl.sflag.p, cl.sscan.p,
                       cl.off.p, cl.max_dev.p);
    GPLUM_HIP(cl.host.ensure(2 * sizeof(int32_t)));
    int32_t *hinfo = cl.host.as<int32_t>();
    GPLUM_HIP(hipMemcpyAsync(hinfo, cl.max_dev.p, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    GPLUM_HIP(hipStreamSynchronize(st));
    const int ncl = hinfo[0];
    cl.ncl = ncl;
    cl.max_size = hinfo[1];
    GPLUM_ENSURE(cl.index_tmp, m);
    hipLaunchKernelGGL(k_order_by_id, dim3(ceil_div(m, 256)), dim3(256), 0, st, m, cl.sflag.p, cl.sscan.p,
                       cl.off.p, cl.index.p, ps.id.p, cl.index_tmp.p);
    cl.index.swap(cl.index_tmp);
    // processing order: costliest cluster first
    GPLUM_ENSURE(cl.size_key, ncl);
    GPLUM_ENSURE(cl.size_key_sorted, ncl);
    GPLUM_ENSURE(cl.order_in, ncl);
    GPLUM_ENSURE(cl.order, ncl);
    hipLaunchKernelGGL(k_cluster_cost, dim3(ceil_div(ncl, 256)), dim3(256), 0, st, ncl, cl.off.p, cl.index.p,
                       ps.dt.p, dt_tree, cl.size_key.p, cl.order_in.p);
    bytes = 0;
    GPLUM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, cl.size_key.p, cl.size_key_sorted.p,
                                                  cl.order_in.p, cl.order.p, ncl, 0, 32, st));
    GPLUM_ENSURE(s.tmp, bytes + 16);
    GPLUM_HIP(hipcub::DeviceRadixSort::SortPairs(s.tmp.p, bytes, cl.size_key.p, cl.size_key_sorted.p,
                                                  cl.order_in.p, cl.order.p, ncl, 0, 32, st));
    GPLUM_HIP(hipGetLastError());
    return GPLUM_OK;
}

}  // namespace gplum
