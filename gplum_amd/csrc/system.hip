// system.hip -- the device-resident P3T system and its C ABI (gplum_system_*).
//
// One soft step follows main_p3t.cpp:423-699 on one GPU:
//   kick -> clusters (union-find over the neighbour graph) -> hard part (batched
//   Hermite on clusters, Kepler drift on isolated particles) -> tree (Morton sort,
//   particles physically reordered, nodes, moments, groups, walk) -> FP32 EP-EP /
//   EP-SP with candidate emission -> FP64 neighbour finalize -> kick -> energy ->
//   (after collisions) merge + full recompute.
// Radii (setCutoffRadii) and the rare merge bookkeeping run on the host in FP64
// with the reference's operation order; everything per-particle per-step is a
// HIP kernel on one stream.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "hermite.hpp"
#include "neighbor.hpp"
#include "soft_kernels.hpp"
#include "system.hpp"
#include "tree.hpp"

namespace gplum {

int ParticleSoA::ensure(int n) {
    GPLUM_ENSURE(id, n); GPLUM_ENSURE(id_cluster, n);
    GPLUM_ENSURE(mass, n); GPLUM_ENSURE(r_planet, n); GPLUM_ENSURE(f, n);
    GPLUM_ENSURE(pos, 3 * n); GPLUM_ENSURE(vel, 3 * n); GPLUM_ENSURE(acc, 3 * n);
    GPLUM_ENSURE(acc_d, 3 * n); GPLUM_ENSURE(jerk_d, 3 * n); GPLUM_ENSURE(acc_s, 3 * n);
    GPLUM_ENSURE(jerk_s, 3 * n);
    GPLUM_ENSURE(phi, n); GPLUM_ENSURE(phi_d, n); GPLUM_ENSURE(phi_s, n); GPLUM_ENSURE(r_out, n);
    GPLUM_ENSURE(r_out_inv, n); GPLUM_ENSURE(r_search, n); GPLUM_ENSURE(v_disp, n);
    GPLUM_ENSURE(time, n); GPLUM_ENSURE(dt, n); GPLUM_ENSURE(acc0, n);
    GPLUM_ENSURE(number, n); GPLUM_ENSURE(rank, n); GPLUM_ENSURE(id_max, n); GPLUM_ENSURE(id_min, n);
    GPLUM_ENSURE(n_cluster, n); GPLUM_ENSURE(flags, n);
    GPLUM_ENSURE(fc_rate, n);
    return GPLUM_OK;
}

// ----------------------------------------------------------------- kernels

// Every particle array at once (one launch instead of one per array): dst[i] = src[perm[i]];
// a null source array is not moved.
struct SoAPtrs {
    double *v3[7];      // pos, vel, acc, acc_d, jerk_d, acc_s, jerk_s
    double *v1[13];     // mass, r_planet, f, phi, phi_d, phi_s, r_out, r_out_inv, r_search, v_disp, time, dt, acc0
    int64_t *l[2];      // id, id_cluster
    int32_t *w[6];      // number, rank, id_max, id_min, n_cluster, flags
    float *f1;          // fc_rate
};

// force_out false: the soft-force outputs (acc, phi, neighbour count, rank, id_max,
// id_min) are left out, for a reorder that a soft force follows (it assigns them for
// every particle).
static SoAPtrs soa_ptrs(ParticleSoA &A, bool force_out = true) {
    return SoAPtrs{{A.pos.p, A.vel.p, force_out ? A.acc.p : nullptr, A.acc_d.p, A.jerk_d.p, A.acc_s.p, A.jerk_s.p},
                   {A.mass.p, A.r_planet.p, A.f.p, force_out ? A.phi.p : nullptr, A.phi_d.p, A.phi_s.p, A.r_out.p,
                    A.r_out_inv.p, A.r_search.p, A.v_disp.p, A.time.p, A.dt.p, A.acc0.p},
                   {A.id.p, A.id_cluster.p},
                   {force_out ? A.number.p : nullptr, force_out ? A.rank.p : nullptr,
                    force_out ? A.id_max.p : nullptr, force_out ? A.id_min.p : nullptr, A.n_cluster.p, A.flags.p},
                   A.fc_rate.p};
}

__global__ void k_gather_all(int n, const int32_t *__restrict__ perm, SoAPtrs src, SoAPtrs dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = perm[i];
#pragma unroll
    for (int k = 0; k < 7; k++)
        if (src.v3[k])
#pragma unroll
            for (int c = 0; c < 3; c++) dst.v3[k][3 * i + c] = src.v3[k][3 * s + c];
#pragma unroll
    for (int k = 0; k < 13; k++)
        if (src.v1[k]) dst.v1[k][i] = src.v1[k][s];
#pragma unroll
    for (int k = 0; k < 2; k++) dst.l[k][i] = src.l[k][s];
#pragma unroll
    for (int k = 0; k < 6; k++)
        if (src.w[k]) dst.w[k][i] = src.w[k][s];
    dst.f1[i] = src.f1[s];
}

__global__ void k_jrec(int n, const double *__restrict__ pos, const double *__restrict__ mass,
                       const double *__restrict__ r_out, const double *__restrict__ r_search, JRec *__restrict__ r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    JRec q;
    q.x = pos[3 * i]; q.y = pos[3 * i + 1]; q.z = pos[3 * i + 2]; q.m = mass[i];
    q.ro = r_out[i]; q.rs = r_search[i]; q.id = i; q.rank = 0;
    r[i] = q;
}

__global__ void k_soft_groups(int ng, const int32_t *__restrict__ gstart, const int32_t *__restrict__ gcount,
                              const int32_t *__restrict__ ep_off, const int32_t *__restrict__ ep_cnt,
                              const int32_t *__restrict__ epf_off, const int32_t *__restrict__ epf_cnt,
                              const int32_t *__restrict__ sp_off, const int32_t *__restrict__ sp_cnt,
                              SoftGroup *__restrict__ out, unsigned long long *__restrict__ pairs) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long ep = 0, epf = 0, sp = 0;
    if (g < ng) {
        SoftGroup s;
        s.i_begin = gstart[g]; s.ni = gcount[g];
        s.j_begin = ep_off[g]; s.nj = ep_cnt[g];
        s.jf_begin = epf_off[g]; s.njf = epf_cnt[g];
        s.s_begin = sp_off[g]; s.ns = sp_cnt[g];
        s.i_ref = gstart[g]; s.pad = 0;
        out[g] = s;
        ep = (unsigned long long)s.ni * (unsigned long long)s.nj;
        epf = (unsigned long long)s.ni * (unsigned long long)s.njf;
        sp = (unsigned long long)s.ni * (unsigned long long)s.ns;
    }
    // interaction counts of the force call (near EP, far EP, SP pairs): a wave sum, then
    // one atomic per wave
    for (int d = 32; d > 0; d >>= 1) {
        ep += __shfl_xor(ep, d);
        epf += __shfl_xor(epf, d);
        sp += __shfl_xor(sp, d);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&pairs[0], ep);
        atomicAdd(&pairs[1], epf);
        atomicAdd(&pairs[2], sp);
    }
}

__global__ void k_kick(int n3, double hdt, const double *__restrict__ acc, double *__restrict__ vel) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    vel[i] += hdt * acc[i];
}

// Energy::calcEnergy partial sums (energy.h:29-63): per block {ekin, esun, ephi, ephid}
__global__ void k_energy(int n, const double *__restrict__ mass, const double *__restrict__ vel,
                         const double *__restrict__ phi, const double *__restrict__ phi_d,
                         const double *__restrict__ phi_s, double *__restrict__ part) {
    __shared__ double s[4][256];
    double e[4] = {0., 0., 0., 0.};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double m = mass[i];
        e[0] += (m * vel[3 * i]) * vel[3 * i] + (m * vel[3 * i + 1]) * vel[3 * i + 1] +
                (m * vel[3 * i + 2]) * vel[3 * i + 2];
        e[1] += m * phi_s[i];
        e[2] += m * phi[i];
        e[3] += m * phi_d[i];
    }
    for (int k = 0; k < 4; k++) s[k][threadIdx.x] = e[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 4; k++) s[k][threadIdx.x] += s[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) part[4 * blockIdx.x + k] = s[k][0];
}

// FPHard(const FPGrav&) for the cluster particles: time_c = time, local time 0
// (particle.h:1096-1116).
__global__ void k_hard_gather(int m, const int32_t *__restrict__ idx, HardPtrs H, PartPtrs P,
                              double *__restrict__ time_c) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int i = idx[k];
    for (int c = 0; c < 3; c++) {
        H.pos[3 * k + c] = P.pos[3 * i + c];
        H.vel[3 * k + c] = P.vel[3 * i + c];
        H.acc_s[3 * k + c] = P.acc_s[3 * i + c];
        H.jerk_s[3 * k + c] = P.jerk_s[3 * i + c];
        H.acc_d[3 * k + c] = P.acc_d[3 * i + c];
        H.jerk_d[3 * k + c] = P.jerk_d[3 * i + c];
    }
    H.mass[k] = P.mass[i];
    H.rinv[k] = P.r_out_inv[i];
    time_c[k] = P.time[i];
    H.time[k] = 0.;
    H.dt[k] = P.dt[i];
    H.acc0[k] = P.acc0[i];
    H.r_planet[k] = P.r_planet[i];
    H.f[k] = P.f[i];
    H.id[k] = P.id[i];
    H.flags[k] = P.flags[i];
}

// FP_t(ptcl_multi[i][j]) after resetTime (hard.h:768-787).
__global__ void k_hard_scatter(int m, const int32_t *__restrict__ idx, const int32_t *__restrict__ off,
                               int ncl, HardPtrs H, PartPtrs P, const double *__restrict__ time_c) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int i = idx[k];
    for (int c = 0; c < 3; c++) {
        P.pos[3 * i + c] = H.pos[3 * k + c];
        P.vel[3 * i + c] = H.vel[3 * k + c];
        P.acc_s[3 * i + c] = H.acc_s[3 * k + c];
        P.jerk_s[3 * i + c] = H.jerk_s[3 * k + c];
        P.acc_d[3 * i + c] = H.acc_d[3 * k + c];
        P.jerk_d[3 * i + c] = H.jerk_d[3 * k + c];
    }
    P.time[i] = time_c[k] + H.time[k];
    P.dt[i] = H.dt[k];
    P.r_planet[i] = H.r_planet[k];
    P.id[i] = H.id[k];
    P.flags[i] = H.flags[k];
    P.phi_d[i] = H.phi_d[k];
    P.phi_s[i] = H.phi_s[k];
    // cluster size: binary search of k in off[0..ncl]
    int lo = 0, hi = ncl;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= k) lo = mid; else hi = mid;
    }
    P.n_cluster[i] = off[lo + 1] - off[lo];
}

// ------------------------------------------------------------ host: radii

// ------------------------------------------------------------------ System

System::System(const gplum_params &prm) : p(prm) {}

System::~System() {
    if (log_rem) std::fclose(log_rem);
    if (log_col) std::fclose(log_col);
    if (st) (void)hipStreamDestroy(st);
    if (st2) (void)hipStreamDestroy(st2);
    if (st_fc) (void)hipStreamSynchronize(st_fc);
    if (st_fc) (void)hipStreamDestroy(st_fc);
    if (st_iso) (void)hipStreamSynchronize(st_iso);
    if (st_iso) (void)hipStreamDestroy(st_iso);
    for (hipEvent_t e : {ev_iso_fork, ev_iso_join})
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {ev_fc_ready, ev_fc_done})
        if (e) (void)hipEventDestroy(e);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    for (auto &e : ev) if (e) (void)hipEventDestroy(e);
    for (auto &e : evd) if (e) (void)hipEventDestroy(e);
    for (const TimedPair &q : tpend) {
        (void)hipEventDestroy(q.a);
        (void)hipEventDestroy(q.b);
    }
    for (hipEvent_t e : ev_free) (void)hipEventDestroy(e);
}

int System::init_device() {
    if (st) return GPLUM_OK;
    int count = 0;
    GPLUM_HIP(hipGetDeviceCount(&count));
    if (count <= 0) return fail(GPLUM_E_NODEVICE, "no HIP device");
    hipDeviceProp_t prop;
    int dev = 0;
    GPLUM_HIP(hipGetDevice(&dev));
    GPLUM_HIP(hipGetDeviceProperties(&prop, dev));
    n_cu = prop.multiProcessorCount;
    GPLUM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    GPLUM_HIP(hipStreamCreateWithFlags(&st2, hipStreamNonBlocking));
    GPLUM_HIP(hipStreamCreateWithFlags(&st_fc, hipStreamNonBlocking));
    GPLUM_HIP(hipStreamCreateWithFlags(&st_iso, hipStreamNonBlocking));
    GPLUM_HIP(hipEventCreateWithFlags(&ev_iso_fork, hipEventDisableTiming));
    GPLUM_HIP(hipEventCreateWithFlags(&ev_iso_join, hipEventDisableTiming));
    GPLUM_HIP(hipEventCreateWithFlags(&ev_fc_ready, hipEventDisableTiming));
    GPLUM_HIP(hipEventCreateWithFlags(&ev_fc_done, hipEventDisableTiming));
    GPLUM_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    GPLUM_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    for (auto &e : ev) GPLUM_HIP(hipEventCreate(&e));
    return GPLUM_OK;
}

PartPtrs System::ptrs(ParticleSoA &s) {
    return PartPtrs{s.pos.p, s.vel.p, s.acc.p, s.acc_d.p, s.jerk_d.p, s.acc_s.p, s.jerk_s.p, s.mass.p,
                    s.r_out.p, s.r_out_inv.p, s.r_search.p, s.time.p, s.dt.p, s.acc0.p, s.r_planet.p,
                    s.f.p, s.phi.p, s.phi_d.p, s.phi_s.p, s.id.p, s.flags.p, s.n_cluster.p, s.number.p};
}

int System::set_particles(const gplum_particle *pp, int count, double time) {
    int rc = init_device();
    if (rc) return rc;
    for (int i = 0; i < count; i++)
        if (pp[i].id < 0 || pp[i].id >= (int64_t(1) << 58))
            return fail(GPLUM_E_ARG, "particle ids must lie in [0, 2^58)");
    n = count;
    time_sys = time;
    if ((rc = A.ensure(n)) || (rc = B.ensure(n))) return rc;
    std::vector<double> v3(3 * (size_t)n), v1(n);
    std::vector<int64_t> i64(n);
    std::vector<int32_t> i32(n);
    auto put3 = [&](DArray<double> &d, size_t off) {
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) v3[3 * i + k] = *(const double *)((const char *)&pp[i] + off + 8 * k);
        return hipMemcpy(d.p, v3.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice);
    };
    auto put1 = [&](DArray<double> &d, size_t off) {
        for (int i = 0; i < n; i++) v1[i] = *(const double *)((const char *)&pp[i] + off);
        return hipMemcpy(d.p, v1.data(), sizeof(double) * n, hipMemcpyHostToDevice);
    };
#define P3(f) GPLUM_HIP(put3(A.f, offsetof(gplum_particle, f)))
#define P1(f) GPLUM_HIP(put1(A.f, offsetof(gplum_particle, f)))
    P3(pos); P3(vel); P3(acc); P3(acc_d); P3(jerk_d); P3(acc_s); P3(jerk_s);
    P1(mass); P1(r_planet); P1(f); P1(phi); P1(phi_d); P1(phi_s); P1(r_out); P1(r_out_inv); P1(r_search);
    P1(v_disp); P1(time); P1(dt); P1(acc0);
#undef P3
#undef P1
    for (int i = 0; i < n; i++) i64[i] = pp[i].id;
    GPLUM_HIP(hipMemcpy(A.id.p, i64.data(), 8 * (size_t)n, hipMemcpyHostToDevice));
    fc.lam_valid = false;
    for (int i = 0; i < n; i++) i64[i] = pp[i].id_cluster;
    GPLUM_HIP(hipMemcpy(A.id_cluster.p, i64.data(), 8 * (size_t)n, hipMemcpyHostToDevice));
    auto puti = [&](DArray<int32_t> &d, auto getter) {
        for (int i = 0; i < n; i++) i32[i] = getter(pp[i]);
        return hipMemcpy(d.p, i32.data(), 4 * (size_t)n, hipMemcpyHostToDevice);
    };
    GPLUM_HIP(puti(A.number, [](const gplum_particle &q) { return q.neighbor.number; }));
    GPLUM_HIP(puti(A.rank, [](const gplum_particle &q) { return q.neighbor.rank; }));
    GPLUM_HIP(puti(A.id_max, [](const gplum_particle &q) { return q.neighbor.id_max; }));
    GPLUM_HIP(puti(A.id_min, [](const gplum_particle &q) { return q.neighbor.id_min; }));
    GPLUM_HIP(puti(A.n_cluster, [](const gplum_particle &q) { return q.n_cluster; }));
    GPLUM_HIP(puti(A.flags, [](const gplum_particle &q) { return q.flags; }));
    nb.n = 0;
    nb.n_nbr = 0;
    have_nbr = false;
    istep = std::llround(time / p.dt_tree);
    return GPLUM_OK;
}

int System::get_particles(gplum_particle *pp, int count) {
    if (count < n) return fail(GPLUM_E_ARG, "output array too small");
    GPLUM_HIP(hipStreamSynchronize(st));
    std::vector<double> v3(3 * (size_t)n), v1(n);
    std::vector<int64_t> i64(n);
    std::vector<int32_t> i32(n);
    std::memset(pp, 0, sizeof(gplum_particle) * (size_t)n);
    auto get3 = [&](DArray<double> &d, size_t off) {
        hipError_t e = hipMemcpy(v3.data(), d.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost);
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) *(double *)((char *)&pp[i] + off + 8 * k) = v3[3 * i + k];
        return e;
    };
    auto get1 = [&](DArray<double> &d, size_t off) {
        hipError_t e = hipMemcpy(v1.data(), d.p, sizeof(double) * n, hipMemcpyDeviceToHost);
        for (int i = 0; i < n; i++) *(double *)((char *)&pp[i] + off) = v1[i];
        return e;
    };
#define G3(f) GPLUM_HIP(get3(A.f, offsetof(gplum_particle, f)))
#define G1(f) GPLUM_HIP(get1(A.f, offsetof(gplum_particle, f)))
    G3(pos); G3(vel); G3(acc); G3(acc_d); G3(jerk_d); G3(acc_s); G3(jerk_s);
    G1(mass); G1(r_planet); G1(f); G1(phi); G1(phi_d); G1(phi_s); G1(r_out); G1(r_out_inv); G1(r_search);
    G1(v_disp); G1(time); G1(dt); G1(acc0);
#undef G3
#undef G1
    GPLUM_HIP(hipMemcpy(i64.data(), A.id.p, 8 * (size_t)n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) pp[i].id = i64[i];
    GPLUM_HIP(hipMemcpy(i64.data(), A.id_cluster.p, 8 * (size_t)n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) pp[i].id_cluster = i64[i];
    auto geti = [&](DArray<int32_t> &d, auto setter) {
        hipError_t e = hipMemcpy(i32.data(), d.p, 4 * (size_t)n, hipMemcpyDeviceToHost);
        for (int i = 0; i < n; i++) setter(pp[i], i32[i]);
        return e;
    };
    GPLUM_HIP(geti(A.number, [](gplum_particle &q, int v) { q.neighbor.number = v; }));
    GPLUM_HIP(geti(A.rank, [](gplum_particle &q, int v) { q.neighbor.rank = v; }));
    GPLUM_HIP(geti(A.id_max, [](gplum_particle &q, int v) { q.neighbor.id_max = v; }));
    GPLUM_HIP(geti(A.id_min, [](gplum_particle &q, int v) { q.neighbor.id_min = v; }));
    GPLUM_HIP(geti(A.n_cluster, [](gplum_particle &q, int v) { q.n_cluster = v; }));
    GPLUM_HIP(geti(A.flags, [](gplum_particle &q, int v) { q.flags = v; }));
    return GPLUM_OK;
}

int System::set_radii() { return device_cutoff_radii(p, A, n, rad, st); }

// Morton sort + physical reorder of the particle arrays (A <- B gathered by perm); the
// soft force that follows assigns its outputs, so they are not moved.
int System::reorder() {
    int rc = tree_sort_keys(tree, A.pos.p, n, st);
    if (rc) return rc;
    return gather_all(tree.perm.p, n, false);
}

// A <- A[perm[0..m)] for every particle array (through B); n becomes m.
int System::gather_all(const int32_t *perm, int m, bool force_out) {
    int rc = B.ensure(std::max(m, 1));
    if (rc) return rc;
    const int gb = ceil_div(m, 256);
    if (m > 0)
        hipLaunchKernelGGL(k_gather_all, dim3(gb), dim3(256), 0, st, m, perm, soa_ptrs(A, force_out), soa_ptrs(B, force_out));
    GPLUM_HIP(hipGetLastError());
    A.swap(B);
    n = m;
    have_nbr = false;   // neighbour indices refer to the old order
    return GPLUM_OK;
}

void System::tick(int k) {
    if (!timing) return;
    if (tuning) {                  // the autotuner reads its intervals at once (elapsed)
        (void)hipEventRecord(ev[k], st);
        return;
    }
    hipEvent_t e = timing_event();
    if (!e) return;
    if (evd[k]) ev_free.push_back(evd[k]);   // an earlier tick that no interval took
    evd[k] = e;
    (void)hipEventRecord(e, st);
}

hipEvent_t System::timing_event() {
    hipEvent_t e = nullptr;
    if (!ev_free.empty()) {
        e = ev_free.back();
        ev_free.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        e = nullptr;
    }
    return e;
}

// *dst += the time between ticks a and b, once both have completed
void System::accum(double *dst, int a, int b) {
    if (!timing) return;
    if (tuning) {
        *dst += elapsed(a, b);
        return;
    }
    if (!evd[a] || !evd[b]) return;
    tpend.push_back(TimedPair{evd[a], evd[b], dst});
    evd[a] = evd[b] = nullptr;
    if (tpend.size() > 1024) resolve_timing();
}

// The pending intervals (their end events are recorded on streams the caller has just
// synchronised, or are waited for here).
void System::resolve_timing() {
    for (const TimedPair &q : tpend) {
        float ms = 0.f;
        if (hipEventSynchronize(q.b) == hipSuccess && hipEventElapsedTime(&ms, q.a, q.b) == hipSuccess) *q.dst += ms;
        ev_free.push_back(q.a);
        ev_free.push_back(q.b);
    }
    tpend.clear();
}

double System::elapsed(int a, int b) {
    float ms = 0.f;
    if (timing && hipEventSynchronize(ev[b]) == hipSuccess && hipEventElapsedTime(&ms, ev[a], ev[b]) == hipSuccess)
        return ms;
    return 0.;
}

int System::fc_join() {
    if (!fc_pending) return GPLUM_OK;
    GPLUM_HIP(hipStreamWaitEvent(st, ev_fc_done, 0));
    fc_pending = false;
    return GPLUM_OK;
}

int System::fc_sync() {
    if (st_fc) GPLUM_HIP(hipStreamSynchronize(st_fc));
    return fc_join();
}

// tree + FP32 soft force + FP64 correction (correctForceLong[Initial])
int System::soft_force(bool initial, bool use_f64) {
    int rc;
    if ((rc = fc_join())) return rc;
    tick(EV_TREE0);
    if ((rc = reorder())) return rc;
    if ((rc = tree_build_nodes(tree, p, A.pos.p, A.mass.p, A.r_search.p, n, st))) return rc;
    if ((rc = tree_walk(tree, p, A.pos.p, A.r_search.p, st))) return rc;
    // the walk has synchronised the stream: the read-backs queued before it are in
    if ((rc = radii_check(rad))) return rc;
    if ((rc = merge_resolve())) return rc;
    tick(EV_TREE1);
    GPLUM_ENSURE(jrec, n);
    hipLaunchKernelGGL(k_jrec, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, A.pos.p, A.mass.p, A.r_out.p,
                       A.r_search.p, jrec.p);
    const int ng = tree.n_group;
    GPLUM_ENSURE(groups, ng);
    GPLUM_ENSURE(pair_cnt, 3);
    GPLUM_HIP(hipMemsetAsync(pair_cnt.p, 0, 3 * sizeof(unsigned long long), st));
    hipLaunchKernelGGL(k_soft_groups, dim3(ceil_div(ng, 256)), dim3(256), 0, st, ng, tree.group_start.p,
                       tree.group_count.p, tree.ep_off.p, tree.ep_cnt.p, tree.epf_off.p, tree.epf_cnt.p,
                       tree.sp_off.p, tree.sp_cnt.p, groups.p, pair_cnt.p);
    SrcRecI is{jrec.p};
    SrcRecJ jn{jrec.p, tree.ep_list.p};
    SrcRecJ jf{jrec.p, tree.epf_list.p};
    SrcNodeSP sp{tree.node_mass.p, tree.node_com.p, tree.node_quad.p, tree.sp_list.p};
    SinkEngine out{A.acc.p, A.phi.p, A.number.p, A.rank.p, A.id_max.p, A.id_min.p};
    GPLUM_ENSURE(cand_count, 1);
    unsigned ncand = 0;
    bool presorted = false;            // nb.keys already holds the sorted candidate keys
    const bool fc_on = fc.enabled && !use_f64;
    // a fresh forecast and training round every predict_every soft forces (not on the
    // recompute after mergers, whose particles have not moved, nor in the autotuner)
    const bool counted = fc_on && !fc.uniform && !skip_train && !tuning;
    if (counted && fc.since_predict < (1 << 30)) fc.since_predict++;
    const bool fc_fresh = fc_on && !fc.uniform && (!fc.lam_valid || (counted && fc.since_predict >= fc.predict_every));
    if (fc_fresh) fc.since_predict = 0;
    const FcGroupView gv{tree.group_start.p, tree.group_count.p, tree.group_node.p, tree.ep_cnt.p,
                         tree.node_bmin.p,   tree.node_bmax.p,    tree.node_rs.p};
    if (fc_on) tick(EV_FCA0);
    if (fc_fresh) {
        if ((rc = fc_features(fc, gv, ng, A.pos.p, A.mass.p, A.v_disp.p, p.m_sun, n, st))) return rc;
    }
    if (fc_on && fc.active()) {
        // learned sizing: one segment per i-group, overflowed groups redone exactly
        if (fc.seg_capacity == 0) fc.seg_capacity = (size_t)last_ncand + last_ncand / 4 + 16 * (size_t)ng + 4096;
        if (fc.uniform) fc.uniform_rate = n > 0 ? (float)((double)last_ncand / n) : 0.f;
        CandBuf cb{nullptr, nullptr, 0u};
        if ((rc = fc_plan(fc, gv, ng, n, A.fc_rate.p, fc_fresh, &cb, st))) return rc;
        tick(EV_FCA1);
        tick(EV_FORCE0);
        if (p.quadrupole)
            launch_soft_f32<SrcRecI, SrcRecJ, SrcNodeSP, SinkEngine, EP | SPQ | EMIT>(groups.p, ng, is, jn, jf, sp,
                                                                                  out, (float)p.eps2, cb, st, soft_tile);
        else
            launch_soft_f32<SrcRecI, SrcRecJ, SrcNodeSP, SinkEngine, EP | SPM | EMIT>(groups.p, ng, is, jn, jf, sp,
                                                                                  out, (float)p.eps2, cb, st, soft_tile);
        tick(EV_FORCE1);
        GPLUM_HIP(hipGetLastError());
        tick(EV_FCB0);
        // the candidates arrive grouped by i-group (ascending i): each group's segment is
        // sorted in LDS straight into the finalize's key array, so the global sort of
        // the other modes is not needed.  The sort is queued before the counts are read
        // back, on a key array sized by the segment storage and the last call's total
        // (sorted again on a larger one in the rare call whose total passes it).
        const int kb = cand_key_bits(n);
        if ((rc = fc_summary(fc, ng, st))) return rc;
        const size_t keys_cap = std::max(fc.seg_capacity, (size_t)last_ncand + last_ncand / 4) + 1;
        GPLUM_ENSURE(nb.keys, keys_cap);
        if ((rc = fc_sort_keys(fc, ng, kb, nb.keys.p, (long long)keys_cap, st))) return rc;
        int64_t info[4] = {0, 0, 0, 0};
        if ((rc = fc_read_info(fc, info, st))) return rc;
        const int64_t requested = info[0], total = info[1], n_ovf = info[2], ovf_total = info[3];
        if (total > (int64_t)INT32_MAX) return fail(GPLUM_E_OVERFLOW, "candidate count exceeds 2^31");
        if ((size_t)total + 1 > keys_cap) {
            GPLUM_ENSURE(nb.keys, (size_t)total + 1);
            if ((rc = fc_sort_keys(fc, ng, kb, nb.keys.p, (long long)total + 1, st))) return rc;
        }
        if (n_ovf > 0) {
            if ((rc = fc_overflow_groups(fc, groups.p, (int)n_ovf, st))) return rc;
            GPLUM_ENSURE(cand, (size_t)ovf_total + 1);
            GPLUM_HIP(hipMemsetAsync(cand_count.p, 0, sizeof(unsigned), st));
            CandBuf cb2{cand.p, cand_count.p, (unsigned)ovf_total};
            launch_soft_f32<SrcRecI, SrcRecJ, SrcNodeSP, SinkDiscard, EP | EMIT | NOFAR>(
                fc.ovf_groups.p, (int)n_ovf, is, jn, jf, sp, SinkDiscard{}, (float)p.eps2, cb2, st, soft_tile);
            unsigned got = 0;
            GPLUM_HIP(hipMemcpyAsync(&got, cand_count.p, sizeof(unsigned), hipMemcpyDeviceToHost, st));
            GPLUM_HIP(hipStreamSynchronize(st));
            if ((int64_t)got != ovf_total)
                return fail(GPLUM_E_STATE, "candidate-only pass found " + std::to_string(got) + " of " +
                                               std::to_string(ovf_total) + " candidates");
            GPLUM_ENSURE(fc.ovf_keys, (size_t)ovf_total + 1);
            if ((rc = cand_sort(cand.p, (int)ovf_total, kb, nb.keys_in, fc.ovf_keys.p, scratch, st))) return rc;
            if ((rc = fc_scatter_overflow(fc, (int)n_ovf, fc.ovf_keys.p, nb.keys.p, st))) return rc;
        }
        tick(EV_FCB1);
        ncand = (unsigned)total;
        presorted = true;
        stats.fc_calls++;
        stats.fc_slots += std::min<int64_t>(requested, (int64_t)fc.seg_capacity);
        stats.fc_ovf_groups += n_ovf;
        stats.fc_ovf_candidates += ovf_total;
        accum(&stats.fc_ms, EV_FCA0, EV_FCA1);
        accum(&stats.fc_ms, EV_FCB0, EV_FCB1);
        // the plan did not fit the segment storage: grow it for the next call
        if ((size_t)requested > fc.seg_capacity) fc.seg_capacity = (size_t)requested + (size_t)requested / 8;
    } else {
        if (fc_on) tick(EV_FCA1);
        // candidate buffer: sized by the previous step's count (or a fraction of n)
        if (cand_cap == 0) cand_cap = std::max<unsigned>(1u << 16, (unsigned)(n / 4));
        for (int attempt = 0; attempt < 3; attempt++) {
            GPLUM_ENSURE(cand, cand_cap);
            GPLUM_HIP(hipMemsetAsync(cand_count.p, 0, sizeof(unsigned), st));
            CandBuf cb{cand.p, cand_count.p, cand_cap};
            tick(EV_FORCE0);
            if (!use_f64) {
                if (p.quadrupole)
                    launch_soft_f32<SrcRecI, SrcRecJ, SrcNodeSP, SinkEngine, EP | SPQ | EMIT>(
                        groups.p, ng, is, jn, jf, sp, out, (float)p.eps2, cb, st, soft_tile);
                else
                    launch_soft_f32<SrcRecI, SrcRecJ, SrcNodeSP, SinkEngine, EP | SPM | EMIT>(
                        groups.p, ng, is, jn, jf, sp, out, (float)p.eps2, cb, st, soft_tile);
            } else {
                if (p.quadrupole)
                    launch_soft_f64<SrcRecI, SrcRecJ, SrcNodeSP, SinkEngine, EP | SPQ | EMIT>(
                        groups.p, ng, is, jn, jf, sp, out, p.eps2, cb, st);
                else
                    launch_soft_f64<SrcRecI, SrcRecJ, SrcNodeSP, SinkEngine, EP | SPM | EMIT>(
                        groups.p, ng, is, jn, jf, sp, out, p.eps2, cb, st);
            }
            tick(EV_FORCE1);
            GPLUM_HIP(hipGetLastError());
            GPLUM_HIP(hipMemcpyAsync(&ncand, cand_count.p, sizeof(unsigned), hipMemcpyDeviceToHost, st));
            GPLUM_HIP(hipStreamSynchronize(st));
            if (ncand <= cand_cap) break;
            cand_cap = ncand + ncand / 4 + 1024;   // too small: grow and redo the step
            stats.cand_retries++;
        }
        if (fc_on) accum(&stats.fc_ms, EV_FCA0, EV_FCA1);
    }
    accum(&stats.force_ms, EV_FORCE0, EV_FORCE1);
    accum(&stats.tree_ms, EV_TREE0, EV_TREE1);
    stats.n_force_calls++;
    if (use_f64) stats.n_force_f64_calls++;
    last_ncand = ncand;
    tick(EV_NB0);
    if ((rc = neighbor_finalize(p, A, n, cand.p, (int)ncand, initial, nb, scratch, st, presorted))) return rc;
    tick(EV_NB1);
    accum(&stats.finalize_ms, EV_NB0, EV_NB1);
    if (fc_fresh && !tuning && !skip_train) {
        // train on this call's exact FP32 candidate counts, on st_fc: the training
        // overlaps what follows (kick, merging, the next hard part)
        GPLUM_HIP(hipEventRecord(ev_fc_ready, st));
        GPLUM_HIP(hipStreamWaitEvent(st_fc, ev_fc_ready, 0));
        hipEvent_t t0 = timing ? timing_event() : nullptr;
        if (t0) GPLUM_HIP(hipEventRecord(t0, st_fc));
        if ((rc = fc_train(fc, nb.number_fp32.p, n, st_fc, false))) return rc;
        hipEvent_t t1 = t0 ? timing_event() : nullptr;
        if (t1) {
            GPLUM_HIP(hipEventRecord(t1, st_fc));
            tpend.push_back(TimedPair{t0, t1, &stats.fc_train_ms});
        } else if (t0) {
            ev_free.push_back(t0);
        }
        GPLUM_HIP(hipEventRecord(ev_fc_done, st_fc));
        fc_pending = true;
    }
    have_nbr = true;
    // the interaction counts stay on the device (pair_cnt) until the statistics are read
    last_groups = ng;
    return GPLUM_OK;
}

int System::energy(double out[5]) {
    const int nb_ = 256;
    GPLUM_ENSURE(eparts, 4 * nb_);
    hipLaunchKernelGGL(k_energy, dim3(nb_), dim3(256), 0, st, n, A.mass.p, A.vel.p, A.phi.p, A.phi_d.p,
                       A.phi_s.p, eparts.p);
    std::vector<double> h(4 * nb_);
    GPLUM_HIP(hipMemcpyAsync(h.data(), eparts.p, sizeof(double) * 4 * nb_, hipMemcpyDeviceToHost, st));
    GPLUM_HIP(hipStreamSynchronize(st));
    double e[4] = {0., 0., 0., 0.};
    for (int b = 0; b < nb_; b++)
        for (int k = 0; k < 4; k++) e[k] += h[4 * b + k];
    out[1] = 0.5 * e[0];
    out[2] = e[1];
    out[3] = 0.5 * e[2];
    out[4] = 0.5 * e[3];
    out[0] = out[1] + out[2] + (out[3] + out[4]);
    return GPLUM_OK;
}

// main_p3t.cpp:289-364: time, flags, r_planet/f defaults, radii, two rounds of
// (tree force + correctForceLongInitial).
int System::initialize() {
    int rc = init_device();
    if (rc) return rc;
    {
        std::vector<gplum_particle> h(n);
        if ((rc = get_particles(h.data(), n))) return rc;
        for (auto &q : h) {
            q.time = time_sys;
            q.neighbor.number = 0;
            q.n_cluster = 0;
            q.flags = 0;
            if (q.r_planet <= 0.) q.r_planet = std::pow(0.75 * q.mass / (3.14159265358979323846 * p.dens), 1. / 3.);
            if (q.f == 0.) q.f = p.increase_factor;
        }
        if ((rc = set_particles(h.data(), n, time_sys))) return rc;
    }
    if ((rc = set_radii())) return rc;
    for (int r = 0; r < 2; r++)
        if ((rc = soft_force(true, plain_f64))) return rc;
    if ((rc = energy(e_init))) return rc;
    return GPLUM_OK;
}

int System::kick() {
    hipLaunchKernelGGL(k_kick, dim3(ceil_div(3 * n, 256)), dim3(256), 0, st, 3 * n, 0.5 * p.dt_tree, A.acc.p,
                       A.vel.p);
    GPLUM_HIP(hipGetLastError());
    return GPLUM_OK;
}

// The per-cluster outcome of the hard part reduced on the device (one workgroup), so
// that the host reads a few words instead of every cluster's entries:
// out[0] block steps, out[1] the largest count of one cluster, out[2] that cluster's size
// (the first such cluster), out[3] pair evaluations, out[4] collisions, out[5] 1 if a
// cluster stopped at the block-step cap; *ed_sum the dissipated energy summed over the
// clusters with a collision in ascending cluster order (the others contribute +0).
constexpr int HSTAT_BLOCK = 1024;
__global__ __launch_bounds__(HSTAT_BLOCK) void k_hard_stats(int ncl, const int *__restrict__ ncol,
                                                            const double *__restrict__ ed,
                                                            const long long *__restrict__ nblock,
                                                            const long long *__restrict__ npair,
                                                            const int32_t *__restrict__ off, int32_t *__restrict__ col,
                                                            long long *__restrict__ out, double *__restrict__ ed_sum) {
    __shared__ unsigned long long s_sum[3];
    __shared__ unsigned long long s_max;
    __shared__ int s_cap, s_base, s_w[HSTAT_BLOCK / 64];
    __shared__ double s_ed[HSTAT_BLOCK];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) { s_sum[0] = s_sum[1] = s_sum[2] = 0ull; s_max = 0ull; s_cap = 0; s_base = 0; }
    __syncthreads();
    unsigned long long bs = 0, pr = 0, nc = 0, mx = 0;
    constexpr int U = 4;              // chunks per round: their loads in flight together
    for (int c00 = 0; c00 < ncl; c00 += U * HSTAT_BLOCK) {
        long long b[U], q[U];
        int h[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int c = c00 + u * HSTAT_BLOCK + t;
            b[u] = c < ncl ? nblock[c] : 0;
            q[u] = c < ncl ? npair[c] : 0;
            h[u] = c < ncl ? ncol[c] : 0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int c = c00 + u * HSTAT_BLOCK + t;
            if (c < ncl) {
                if (b[u] < 0) s_cap = 1;
                else {
                    bs += (unsigned long long)b[u];
                    // largest count first, then the lowest cluster index
                    const unsigned long long key = ((unsigned long long)b[u] << 32) | (unsigned)(0x7fffffff - c);
                    mx = key > mx ? key : mx;
                }
                pr += (unsigned long long)q[u];
                nc += (unsigned long long)h[u];
            }
            const int hit = (c < ncl && h[u] > 0) ? 1 : 0;
            // clusters with a collision, listed in ascending index
            const unsigned long long m = __ballot(hit);
            if (lane == 0) s_w[w] = __popcll(m);
            __syncthreads();
            int before = s_base;
            for (int k = 0; k < w; k++) before += s_w[k];
            if (hit) col[before + __popcll(m & ((1ull << lane) - 1ull))] = c;
            __syncthreads();
            if (t == 0) for (int k = 0; k < HSTAT_BLOCK / 64; k++) s_base += s_w[k];
            __syncthreads();
        }
    }
    atomicAdd(&s_sum[0], bs);
    atomicAdd(&s_sum[1], pr);
    atomicAdd(&s_sum[2], nc);
    atomicMax(&s_max, mx);
    __syncthreads();
    // the dissipated energies of the collided clusters, gathered by all threads and
    // summed by one in ascending cluster order
    double e = 0.;
    for (int k0 = 0; k0 < s_base; k0 += HSTAT_BLOCK) {
        if (k0 + t < s_base) s_ed[t] = ed[col[k0 + t]];
        __syncthreads();
        if (t == 0) {
            const int kn = min(HSTAT_BLOCK, s_base - k0);
            for (int k = 0; k < kn; k++) e += s_ed[k];
        }
        __syncthreads();
    }
    if (t == 0) {
        const int cm = 0x7fffffff - (int)(unsigned)(s_max & 0xffffffffull);
        out[0] = (long long)s_sum[0];
        out[1] = (long long)(s_max >> 32);
        out[2] = (s_max != 0ull && cm >= 0 && cm < ncl) ? (long long)(off[cm + 1] - off[cm]) : 0;
        out[3] = (long long)s_sum[1];
        out[4] = (long long)s_sum[2];
        out[5] = s_cap;
        *ed_sum = e;
    }
}

// hard part: clusters + isolated particles (HardSystem::timeIntegrate, hard.h:554-829)
int System::hard_part(bool persistent) {
    int rc;
    if (!have_nbr) return fail(GPLUM_E_STATE, "neighbour graph missing: initialize first");
    tick(EV_HARD0);
    // the isolated particles (Kepler drift) on their own stream, alongside the cluster
    // search and the clusters: the two sets are disjoint (number == 0 vs > 0), and the
    // search reads only the neighbour graph, ids and the members' steps
    IsoArgs ia{A.pos.p, A.vel.p, A.acc_d.p, A.jerk_d.p, A.acc_s.p, A.jerk_s.p, A.phi_d.p, A.phi_s.p, A.time.p,
               A.dt.p, A.acc0.p, A.number.p, A.n_cluster.p};
    GPLUM_HIP(hipEventRecord(ev_iso_fork, st));
    GPLUM_HIP(hipStreamWaitEvent(st_iso, ev_iso_fork, 0));
    if ((rc = isolated_launch(p, ia, n, time_sys, time_sys + p.dt_tree, st_iso))) return rc;
    GPLUM_HIP(hipEventRecord(ev_iso_join, st_iso));
    if ((rc = make_clusters(A, n, nb, cl, scratch, p.dt_tree, st))) {
        (void)hipStreamSynchronize(st_iso);
        return rc;
    }
    const int m = cl.m, ncl = cl.ncl;
    stats.n_clusters = ncl;
    stats.n_cluster_particles = m;
    stats.max_cluster = cl.max_size;
    step_ncol = 0;
    step_edisp = 0.;
    if (ncl > 0) {
        if ((rc = hb.ensure(m))) return rc;
        hb.max_size = cl.max_size;
        GPLUM_ENSURE(time_c, m);
        HardPtrs H = hard_ptrs();
        PartPtrs P = ptrs(A);
        hipLaunchKernelGGL(k_hard_gather, dim3(ceil_div(m, 256)), dim3(256), 0, st, m, cl.index.p, H, P, time_c.p);
        // collision records for the collision log (at most one collision per particle)
        hb.crec_cap = 0;
        if (log_col) {
            const unsigned cap = (unsigned)std::min<long long>((long long)m / 2 + 16, 1ll << 20);
            GPLUM_ENSURE(hb.crec, (size_t)cap * COLREC);
            GPLUM_ENSURE(hb.ccount, 1);
            GPLUM_HIP(hipMemsetAsync(hb.ccount.p, 0, sizeof(unsigned), st));
            hb.crec_cap = cap;
        }
        if (persistent) {
            tick(EV_HK0);
            if ((rc = hard_launch(p, hb, ncl, cl.off.p, cl.order.p, true, n_cu, st, st2, ev_fork, ev_join,
                                  hard_wg_per_cu, hard_lds_max, hard_giant_min, hard_terms_max)))
                return rc;
            tick(EV_HK1);
        } else {
            // the comparator: one launch per cluster
            GPLUM_ENSURE(hb.next, 3);
            GPLUM_ENSURE(hb.edisp, ncl);
            GPLUM_ENSURE(hb.ncol, ncl);
            GPLUM_ENSURE(hb.nblock, ncl);
            GPLUM_ENSURE(hb.npair, ncl);
            std::vector<int32_t> hoff(ncl + 1);
            GPLUM_HIP(hipMemcpyAsync(hoff.data(), cl.off.p, sizeof(int32_t) * (ncl + 1), hipMemcpyDeviceToHost, st));
            GPLUM_HIP(hipStreamSynchronize(st));
            for (int c = 0; c < ncl; c++)
                if ((rc = hard_launch_single(p, hb, c, hoff[c + 1] - hoff[c], cl.off.p, st))) return rc;
        }
        hipLaunchKernelGGL(k_hard_scatter, dim3(ceil_div(m, 256)), dim3(256), 0, st, m, cl.index.p, cl.off.p, ncl,
                           H, P, time_c.p);
        GPLUM_ENSURE(hstat, 8);
        GPLUM_ENSURE(hstat_col, ncl + 1);
        hipLaunchKernelGGL(k_hard_stats, dim3(1), dim3(HSTAT_BLOCK), 0, st, ncl, hb.ncol.p, hb.edisp.p, hb.nblock.p,
                           hb.npair.p, cl.off.p, hstat_col.p, hstat.p, reinterpret_cast<double *>(hstat.p + 6));
        GPLUM_HIP(hst.ensure(sizeof(long long) * 8));
        long long *hs = hst.as<long long>();
        GPLUM_HIP(hipMemcpyAsync(hs, hstat.p, sizeof(long long) * 7, hipMemcpyDeviceToHost, st));
        GPLUM_HIP(hipStreamSynchronize(st));
        resolve_timing();          // the earlier intervals have completed with the stream
        if (hs[5]) return fail(GPLUM_E_STATE, "hard cluster exceeded the block-step cap");
        double ed_sum;
        std::memcpy(&ed_sum, &hs[6], sizeof(double));
        stats.block_steps += hs[0];
        if (hs[1] > stats.max_block_steps) {
            stats.max_block_steps = hs[1];
            stats.max_block_cluster = (int)hs[2];
        }
        stats.hard_pairs += hs[3];
        step_ncol = (int)hs[4];
        step_edisp = ed_sum;
        if (log_col && step_ncol > 0 && (rc = write_collision_log())) return rc;
    }
    GPLUM_HIP(hipStreamWaitEvent(st, ev_iso_join, 0));
    tick(EV_HARD1);
    accum(&stats.hard_ms, EV_HARD0, EV_HARD1);
    if (ncl > 0 && persistent) {
        accum(&stats.hard_kernel_ms, EV_HK0, EV_HK1);
        stats.n_hard_launches++;
    }
    return GPLUM_OK;
}

// The collision log (Collision0::write2File, collisionA.h:122-148; written by rank 0 in
// the order of the collision list, hard.h:978): one line per merger of the last hard
// part, clusters in processing order and each cluster's mergers in time order.  Perfect
// merging: no fragments (n_frag 0, id_frag 0, mass_frag 0) and Flag = flag_merge = 1.
int System::write_collision_log() {
    GPLUM_HIP(hst.ensure(sizeof(unsigned)));
    GPLUM_HIP(hipMemcpyAsync(hst.p, hb.ccount.p, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    GPLUM_HIP(hipStreamSynchronize(st));
    const unsigned k = *hst.as<unsigned>();
    if (k > hb.crec_cap) return fail(GPLUM_E_OVERFLOW, "more collisions than collision-log records");
    if (k == 0) return GPLUM_OK;
    GPLUM_HIP(hst.ensure(sizeof(double) * COLREC * (size_t)k));
    const double *r = hst.as<double>();
    GPLUM_HIP(hipMemcpyAsync(hst.p, hb.crec.p, sizeof(double) * COLREC * (size_t)k, hipMemcpyDeviceToHost, st));
    GPLUM_HIP(hipStreamSynchronize(st));
    std::vector<unsigned> ord(k);
    for (unsigned a = 0; a < k; a++) ord[a] = a;
    std::sort(ord.begin(), ord.end(), [&](unsigned x, unsigned y) {
        const double *u = r + (size_t)x * COLREC, *v = r + (size_t)y * COLREC;
        if (u[COLREC - 1] != v[COLREC - 1]) return u[COLREC - 1] < v[COLREC - 1];
        return u[0] < v[0];
    });
    for (unsigned a : ord) {
        const double *q = r + (size_t)a * COLREC;
        std::fprintf(log_col, "%.8f\t%lld\t%lld\t0\t0\t%.15e\t%.15e\t%.15e", time_sys + q[0], (long long)q[1],
                     (long long)q[2], q[3], q[4], 0.);
        for (int c = 5; c < 17; c++) std::fprintf(log_col, "\t%.15e", q[c]);
        std::fprintf(log_col, "\t1\t%.15e\t%.15e\n", q[17], q[18]);
    }
    std::fflush(log_col);
    return GPLUM_OK;
}

HardPtrs System::hard_ptrs() {
    return HardPtrs{hb.pos.p, hb.vel.p, hb.acc_s.p, hb.jerk_s.p, hb.acc_d.p, hb.jerk_d.p, hb.mass.p, hb.rinv.p,
                    hb.time.p, hb.dt.p, hb.acc0.p, hb.r_planet.p, hb.f.p, hb.phi_d.p, hb.phi_s.p, hb.id.p,
                    hb.flags.p};
}

// MergeParticle (func.h:160-207) on the host, then setCutoffRadii + force +
// correctForceLongInitial (main_p3t.cpp:640-699).
// MergeParticle (func.h:160-207) on the device, then the recomputation of
// main_p3t.cpp:640-699 (setCutoffRadii, tree force, correctForceLongInitial).  After
// the hard part a merged body is its target (MERGED) plus dead twins carrying the
// target's id; only those few particles are listed on the host to pair them up.
__global__ void k_merge_flag(int n, const int32_t *__restrict__ flags, int32_t *__restrict__ sel,
                             int32_t *__restrict__ keep) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int f = flags[i];
    sel[i] = (f & (GPLUM_FLAG_DEAD | GPLUM_FLAG_MERGED)) ? 1 : 0;
    keep[i] = (f & GPLUM_FLAG_DEAD) ? 0 : 1;
}

__global__ void k_compact_idx(int n, const int32_t *__restrict__ flag, const int32_t *__restrict__ scan,
                              int32_t *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) out[scan[i]] = i;
}

// the flags and ids of the listed particles (the merged bodies found by k_post_flags)
__global__ void k_sel_info(int k, const int32_t *__restrict__ sel, const int32_t *__restrict__ flags,
                           const int64_t *__restrict__ id, int32_t *__restrict__ fl_out, int64_t *__restrict__ id_out) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= k) return;
    fl_out[a] = flags[sel[a]];
    id_out[a] = id[sel[a]];
}

// one thread per target: its twins in ascending index, the reference's arithmetic
__global__ void k_merge_apply(int ntg, const int32_t *__restrict__ tgt, const int32_t *__restrict__ twin_off,
                              const int32_t *__restrict__ twin, double *__restrict__ mass, double *__restrict__ vel,
                              double *__restrict__ phi, double *__restrict__ phi_d, int32_t *__restrict__ flags,
                              double *__restrict__ ed) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ntg) return;
    const int i = tgt[k];
    double e = 0.;
    for (int a = twin_off[k]; a < twin_off[k + 1]; a++) {
        const int j = twin[a];
        const double mi = mass[i], mj = mass[j];
        const double vr[3] = {vel[3 * j] - vel[3 * i], vel[3 * j + 1] - vel[3 * i + 1], vel[3 * j + 2] - vel[3 * i + 2]};
        mass[i] += mj;
        for (int c = 0; c < 3; c++) vel[3 * i + c] = (mi * vel[3 * i + c] + mj * vel[3 * j + c]) / (mi + mj);
        phi[i] = (mi * phi[i] + mj * phi[j]) / (mi + mj);
        phi_d[i] = (mi * phi_d[i] + mj * phi_d[j]) / (mi + mj);
        const double s = 0.5 * mi * mj / (mi + mj);   // `0.5*mi*mj/(mi+mj) * vrel*vrel`: (s v).v
        e -= (s * vr[0]) * vr[0] + (s * vr[1]) * vr[1] + (s * vr[2]) * vr[2];
    }
    flags[i] &= ~GPLUM_FLAG_MERGED;
    ed[k] = e;
}

template <class T>
static int excl_sum(DArray<unsigned char> &tmp, const T *in, T *out, int n, hipStream_t st) {
    size_t bytes = 0;
    GPLUM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, st));
    GPLUM_ENSURE(tmp, bytes + 16);
    GPLUM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, n, st));
    return GPLUM_OK;
}

// Flags after the step's force: sel = DEAD or MERGED (the twins and targets of this
// step's mergers), rem = live and outside [r_min, r_max] (removeParticlesOutOfBoundary,
// func.h:237-247: |x|^2 > r_max^2 or |x|^2 < r_min^2), keep = live and inside.  Merging
// moves no live particle, so the bounds can be tested before it.  cnt[0] counts sel,
// cnt[1] rem.
__global__ void k_post_flags(int n, const double *__restrict__ pos, const int32_t *__restrict__ flags, double rmax2,
                             double rmin2, int32_t *__restrict__ sel, int32_t *__restrict__ keep,
                             int32_t *__restrict__ rem, int32_t *__restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int f = flags[i];
    const double pos2 = pos[3 * i] * pos[3 * i] + pos[3 * i + 1] * pos[3 * i + 1] + pos[3 * i + 2] * pos[3 * i + 2];
    const bool live = !(f & GPLUM_FLAG_DEAD);
    const bool out = live && (pos2 > rmax2 || pos2 < rmin2);
    const int s = (f & (GPLUM_FLAG_DEAD | GPLUM_FLAG_MERGED)) ? 1 : 0;
    sel[i] = s;
    rem[i] = out ? 1 : 0;
    keep[i] = (live && !out) ? 1 : 0;
    if (s) atomicAdd(&cnt[0], 1);
    if (out) atomicAdd(&cnt[1], 1);
    if (!(live && !out)) atomicAdd(&cnt[2], 1);     // dropped by the compaction
}

// the removed particles' fields for the energy bookkeeping and the remove log
__global__ void k_gather_removed(int k, const int32_t *__restrict__ idx, PartPtrs P, double *__restrict__ out,
                                 int64_t *__restrict__ ids) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= k) return;
    const int i = idx[a];
    double *o = out + 11 * (size_t)a;
    o[0] = P.time[i];
    o[1] = P.mass[i];
    for (int c = 0; c < 3; c++) { o[2 + c] = P.pos[3 * i + c]; o[5 + c] = P.vel[3 * i + c]; }
    o[8] = P.phi[i];
    o[9] = P.phi_d[i];
    o[10] = P.phi_s[i];
    ids[a] = P.id[i];
}

// MergeParticle (func.h:160-207) and removeParticlesOutOfBoundary (func.h:209-380) on the
// device, one compaction for both.  After the hard part a merged body is its target
// (MERGED) plus dead twins carrying the target's id; only those few particles are listed
// on the host to pair them up.  The removed particles (normally none, a handful at
// most) are copied to the host, which books their energy as the reference's rank 0 does
// (func.h:305-333, taken in ascending id) and writes the remove log (func.h:347-352).
// The merged targets' energies of the last merge_and_remove, added to edisp in target
// order (the caller has synchronised the stream since).
int System::merge_resolve() {
    MergeScratch &M = mrg;
    if (M.ed_pending > 0) {
        const double *ed = M.ed_host.as<double>();
        for (int a = 0; a < M.ed_pending; a++) edisp += ed[a];
        M.ed_pending = 0;
    }
    return GPLUM_OK;
}

int System::merge_and_remove(int *n_removed) {
    *n_removed = 0;
    MergeScratch &M = mrg;
    GPLUM_ENSURE(M.flag, n + 1);
    GPLUM_ENSURE(M.scan, n + 1);
    GPLUM_ENSURE(M.keep, n + 1);
    GPLUM_ENSURE(M.kscan, n + 1);
    GPLUM_ENSURE(M.rem, n + 1);
    GPLUM_ENSURE(M.rscan, n + 1);
    GPLUM_ENSURE(M.count, 3);
    GPLUM_HIP(hipMemsetAsync(M.count.p, 0, 3 * sizeof(int32_t), st));
    hipLaunchKernelGGL(k_post_flags, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, A.pos.p, A.flags.p,
                       r_max * r_max, r_min * r_min, M.flag.p, M.keep.p, M.rem.p, M.count.p);
    GPLUM_HIP(hst.ensure(sizeof(int32_t) * 3));
    GPLUM_HIP(hipMemcpyAsync(hst.p, M.count.p, sizeof(int32_t) * 3, hipMemcpyDeviceToHost, st));
    GPLUM_HIP(hipStreamSynchronize(st));
    const int k = hst.as<int32_t>()[0], n_rem = hst.as<int32_t>()[1], n_drop = hst.as<int32_t>()[2];
    if (k == 0 && n_rem == 0) return GPLUM_OK;
    int rc;
    if (k > 0) {
        if ((rc = excl_sum(rad.tmp, M.flag.p, M.scan.p, n, st))) return rc;
        GPLUM_ENSURE(M.sel, k);
        hipLaunchKernelGGL(k_compact_idx, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, M.flag.p, M.scan.p, M.sel.p);
        // the listed particles' index, flags and id in one read
        GPLUM_ENSURE(M.sel_info, 2 * (size_t)k);
        int32_t *d_fl = reinterpret_cast<int32_t *>(M.sel_info.p);
        int64_t *d_id = M.sel_info.p + k;
        hipLaunchKernelGGL(k_sel_info, dim3(ceil_div(k, 64)), dim3(64), 0, st, k, M.sel.p, A.flags.p, A.id.p, d_fl,
                           d_id);
        GPLUM_HIP(hst.ensure(16 * (size_t)k));
        int64_t *h_id = hst.as<int64_t>();
        int32_t *h_sel = reinterpret_cast<int32_t *>(h_id + k);
        int32_t *h_fl = h_sel + k;
        GPLUM_HIP(hipMemcpyAsync(h_id, d_id, sizeof(int64_t) * k, hipMemcpyDeviceToHost, st));
        GPLUM_HIP(hipMemcpyAsync(h_sel, M.sel.p, sizeof(int32_t) * k, hipMemcpyDeviceToHost, st));
        GPLUM_HIP(hipMemcpyAsync(h_fl, d_fl, sizeof(int32_t) * k, hipMemcpyDeviceToHost, st));
        GPLUM_HIP(hipStreamSynchronize(st));
        const std::vector<int32_t> sel(h_sel, h_sel + k), fl(h_fl, h_fl + k);
        const std::vector<int64_t> ids(h_id, h_id + k);
        // pair twins with their target (same id), both in ascending index
        std::vector<int32_t> tgt, off{0}, twin;
        for (int a = 0; a < k; a++) {
            if ((fl[a] & GPLUM_FLAG_DEAD) || !(fl[a] & GPLUM_FLAG_MERGED)) continue;
            tgt.push_back(sel[a]);
            for (int b = 0; b < k; b++)
                if ((fl[b] & GPLUM_FLAG_DEAD) && ids[b] == ids[a]) twin.push_back(sel[b]);
            off.push_back((int32_t)twin.size());
        }
        int n_dead = 0;
        for (int a = 0; a < k; a++) n_dead += (fl[a] & GPLUM_FLAG_DEAD) ? 1 : 0;
        if (n_dead != (int)twin.size()) return fail(GPLUM_E_STATE, "a merged particle has no live target");
        const int ntg = (int)tgt.size();
        if (ntg > 0) {
            GPLUM_ENSURE(M.tgt, ntg);
            GPLUM_ENSURE(M.twin_off, ntg + 1);
            GPLUM_ENSURE(M.twin, twin.size() + 1);
            GPLUM_ENSURE(M.ed, ntg);
            GPLUM_HIP(hipMemcpyAsync(M.tgt.p, tgt.data(), sizeof(int32_t) * ntg, hipMemcpyHostToDevice, st));
            GPLUM_HIP(hipMemcpyAsync(M.twin_off.p, off.data(), sizeof(int32_t) * (ntg + 1), hipMemcpyHostToDevice, st));
            if (!twin.empty())
                GPLUM_HIP(hipMemcpyAsync(M.twin.p, twin.data(), sizeof(int32_t) * twin.size(), hipMemcpyHostToDevice,
                                         st));
            hipLaunchKernelGGL(k_merge_apply, dim3(ceil_div(ntg, 64)), dim3(64), 0, st, ntg, M.tgt.p, M.twin_off.p,
                               M.twin.p, A.mass.p, A.vel.p, A.phi.p, A.phi_d.p, A.flags.p, M.ed.p);
            // the merged targets' energies are read back without a wait and added to
            // edisp, in target order, at the next synchronisation (merge_resolve)
            GPLUM_HIP(M.ed_host.ensure(sizeof(double) * ntg));
            GPLUM_HIP(hipMemcpyAsync(M.ed_host.p, M.ed.p, sizeof(double) * ntg, hipMemcpyDeviceToHost, st));
            M.ed_pending = ntg;
        }
    }
    if (n_rem > 0) {
        GPLUM_HIP(hipStreamSynchronize(st));
        if ((rc = merge_resolve())) return rc;    // the mergers' energies before the removals'
        // the removed particles after the merge (a target's mass, velocity and potentials
        // are its merged ones), ascending index, then ascending id
        if ((rc = excl_sum(rad.tmp, M.rem.p, M.rscan.p, n, st))) return rc;
        GPLUM_ENSURE(M.ridx, n_rem);
        GPLUM_ENSURE(M.rdata, 11 * (size_t)n_rem);
        GPLUM_ENSURE(M.rid, n_rem);
        hipLaunchKernelGGL(k_compact_idx, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, M.rem.p, M.rscan.p, M.ridx.p);
        hipLaunchKernelGGL(k_gather_removed, dim3(ceil_div(n_rem, 64)), dim3(64), 0, st, n_rem, M.ridx.p, ptrs(A),
                           M.rdata.p, M.rid.p);
        std::vector<double> rd(11 * (size_t)n_rem);
        std::vector<int64_t> rid(n_rem);
        GPLUM_HIP(hipMemcpyAsync(rd.data(), M.rdata.p, sizeof(double) * rd.size(), hipMemcpyDeviceToHost, st));
        GPLUM_HIP(hipMemcpyAsync(rid.data(), M.rid.p, sizeof(int64_t) * n_rem, hipMemcpyDeviceToHost, st));
        GPLUM_HIP(hipStreamSynchronize(st));
        std::vector<Removed> R(n_rem);
        for (int a = 0; a < n_rem; a++) {
            const double *o = &rd[11 * (size_t)a];
            Removed &q = R[a];
            q.id = rid[a];
            q.time = o[0];
            q.mass = o[1];
            for (int c = 0; c < 3; c++) { q.pos[c] = o[2 + c]; q.vel[c] = o[5 + c]; }
            q.phi = o[8];
            q.phi_d = o[9];
            q.phi_s = o[10];
        }
        std::stable_sort(R.begin(), R.end(), [](const Removed &x, const Removed &y) { return x.id < y.id; });
        double loc = 0.;
        for (int a = 0; a < n_rem; a++) {
            const Removed &q = R[a];
            const double mi = q.mass, s = 0.5 * mi;
            loc -= (s * q.vel[0]) * q.vel[0] + (s * q.vel[1]) * q.vel[1] + (s * q.vel[2]) * q.vel[2];
            loc -= mi * q.phi_s;
            loc -= mi * q.phi_d;
            loc -= mi * q.phi;
            for (int b = 0; b < a; b++) {
                const Removed &r = R[b];
                if (q.id == r.id) continue;
                const double dr[3] = {q.pos[0] - r.pos[0], q.pos[1] - r.pos[1], q.pos[2] - r.pos[2]};
                const double rinv = 1. / std::sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2] + p.eps2);
                loc += -mi * r.mass * rinv;
            }
            if (log_rem)
                std::fprintf(log_rem, "%.8f\t%lld\t%.15e\t%.15e\t%.15e\t%.15e\t%.15e\t%.15e\t%.15e\n", q.time,
                             (long long)q.id, q.mass, q.pos[0], q.pos[1], q.pos[2], q.vel[0], q.vel[1], q.vel[2]);
        }
        if (log_rem) std::fflush(log_rem);
        edisp += loc;
        n_remove_tot += n_rem;
        *n_removed = n_rem;
    }
    // removeParticle: keep the live particles inside the boundary in their order (their
    // count known from the flags' read-back)
    if ((rc = excl_sum(rad.tmp, M.keep.p, M.kscan.p, n, st))) return rc;
    const int m = n - n_drop;
    GPLUM_ENSURE(M.perm, m + 1);
    hipLaunchKernelGGL(k_compact_idx, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, M.keep.p, M.kscan.p, M.perm.p);
    // a merge or a removal is always followed by the recomputed soft force, which assigns
    // the force outputs of every particle: they are not moved
    return gather_all(M.perm.p, m, false);
}

// Energy::calcEnergy of the step (main_p3t.cpp:629, 692): the partial sums stay on the
// device; gplum_system_energy reduces the state it is asked about.
int System::energy_launch() {
    GPLUM_ENSURE(eparts_step, 4 * 256);
    hipLaunchKernelGGL(k_energy, dim3(256), dim3(256), 0, st, n, A.mass.p, A.vel.p, A.phi.p, A.phi_d.p, A.phi_s.p,
                       eparts_step.p);
    GPLUM_HIP(hipGetLastError());
    return GPLUM_OK;
}

// One timed run of the batched Hermite kernels on the current clusters (gathered from
// the system, integrated in the gathered copy, not scattered back): the system's state
// is unchanged.
int System::hard_trial(int wg_per_cu, int lds_max, double *ms) {
    int rc;
    *ms = 0.;
    if ((rc = make_clusters(A, n, nb, cl, scratch, p.dt_tree, st))) return rc;
    const int m = cl.m, ncl = cl.ncl;
    if (ncl == 0) return GPLUM_OK;
    if ((rc = hb.ensure(m))) return rc;
    hb.max_size = cl.max_size;
    GPLUM_ENSURE(time_c, m);
    HardPtrs H = hard_ptrs();
    PartPtrs P = ptrs(A);
    hipLaunchKernelGGL(k_hard_gather, dim3(ceil_div(m, 256)), dim3(256), 0, st, m, cl.index.p, H, P, time_c.p);
    tick(EV_HARD0);
    if ((rc = hard_launch(p, hb, ncl, cl.off.p, cl.order.p, true, n_cu, st, st2, ev_fork, ev_join, wg_per_cu,
                          lds_max, hard_giant_min, hard_terms_max)))
        return rc;
    tick(EV_HARD1);
    *ms = elapsed(EV_HARD0, EV_HARD1);
    GPLUM_HIP(hipStreamSynchronize(st));
    return GPLUM_OK;
}

// The project's autotuner: times the soft stage (tree, FP32 force, neighbour finalize)
// for each i-group size x j-tile width and the Hermite kernels for each persistent grid
// x LDS binning threshold on the system's current state, best of `trials` runs each,
// and keeps the fastest pair of each stage.  The i-group size changes which tree nodes
// are opened (the tree approximation, as the reference's n_group_limit does) and the
// tile width the FP32 partial sums; neighbour lists are decided in FP64 and do not
// change, and the FP32 force stays within its tolerance of the FP64 force on the same
// lists.  The hard knobs change no result.  The soft force is recomputed with
// the chosen shape at the end, so the state is that of a soft force with the choice.
int System::autotune(int trials, gplum_tuning *out) {
    if (!have_nbr) return fail(GPLUM_E_STATE, "autotune needs an initialised system");
    trials = std::max(1, trials);
    const gplum_stats saved = stats;
    const bool timing_saved = timing;
    timing = true;
    tuning = true;
    int rc = GPLUM_OK;
    const int groups_c[3] = {64, 128, 256}, tiles_c[3] = {64, 128, 256};
    const int wg_c[4] = {2, 4, 8, 16}, lds_c[2] = {128, 256};
    double best_soft = 1e300, best_hard = 1e300;
    int b_group = p.n_group_limit, b_tile = soft_tile, b_wg = hard_wg_per_cu, b_lds = hard_lds_max;
    for (int gi = 0; gi < 3 && !rc; gi++)
        for (int ti = 0; ti < 3 && !rc; ti++) {
            p.n_group_limit = groups_c[gi];
            soft_tile = tiles_c[ti];
            double t = 1e300;
            for (int r = 0; r < trials && !rc; r++) {
                if ((rc = soft_force(false, false))) break;
                t = std::min(t, elapsed(EV_TREE0, EV_NB1));
            }
            if (!rc && t < best_soft) { best_soft = t; b_group = groups_c[gi]; b_tile = tiles_c[ti]; }
        }
    for (int wi = 0; wi < 4 && !rc; wi++)
        for (int li = 0; li < 2 && !rc; li++) {
            double t = 1e300;
            for (int r = 0; r < trials && !rc; r++) {
                double ms = 0.;
                if ((rc = hard_trial(wg_c[wi], lds_c[li], &ms))) break;
                t = std::min(t, ms);
            }
            if (!rc && t < best_hard) { best_hard = t; b_wg = wg_c[wi]; b_lds = lds_c[li]; }
        }
    p.n_group_limit = b_group;
    soft_tile = b_tile;
    hard_wg_per_cu = b_wg;
    hard_lds_max = b_lds;
    if (!rc) rc = soft_force(false, false);
    tuning = false;
    timing = timing_saved;
    stats = saved;
    if (rc) return rc;
    if (out) {
        out->n_group_limit = b_group;
        out->soft_tile = b_tile;
        out->hard_wg_per_cu = b_wg;
        out->hard_lds_max = b_lds;
        out->hard_giant_min = hard_giant_min;
        out->tree_coop_min = tree.coop_min;
        out->hard_terms_max = hard_terms_max;
        out->soft_ms = best_soft;
        out->hard_ms = best_hard < 1e300 ? best_hard : 0.;
    }
    return GPLUM_OK;
}

// One soft step in the order of main_p3t.cpp:423-699: kick, hard part, soft force +
// correctForceLong, kick, energy, MergeParticle, removeParticlesOutOfBoundary and, after
// a collision, a removal or every reset_step-th step, setCutoffRadii + soft force +
// correctForceLongInitial + energy (the domain decomposition of that branch is a no-op on
// one GPU).
int System::step(bool persistent) {
    int rc;
    if ((rc = kick())) return rc;
    if ((rc = hard_part(persistent && !plain_per_cluster))) return rc;
    n_col_tot += step_ncol;
    edisp += step_edisp;
    if ((rc = soft_force(false, plain_f64))) return rc;
    if ((rc = kick())) return rc;
    if ((rc = energy_launch())) return rc;
    int n_removed = 0;
    if ((rc = merge_and_remove(&n_removed))) return rc;
    const bool reset = reset_step > 0 && istep % reset_step == reset_step - 1;
    if (step_ncol || n_removed || reset) {
        if ((rc = set_radii())) return rc;
        // the recomputed force repeats the step's neighbour counts for all but the merged
        // or removed bodies: the forecaster plans with it but does not train on it again
        skip_train = true;
        rc = soft_force(true, plain_f64);
        skip_train = false;
        if (rc) return rc;
        if ((rc = energy_launch())) return rc;
        n_recompute++;
    }
    time_sys += p.dt_tree;
    istep++;
    stats.steps++;
    return GPLUM_OK;
}

}  // namespace gplum

// ------------------------------------------------------------------ C ABI

using gplum::System;

extern "C" {

int gplum_system_create(const gplum_params *p, gplum_system **out) {
    if (!p || !out) return gplum::fail(GPLUM_E_ARG, "null argument");
    System *s = new System(*p);
    int rc = s->init_device();
    if (rc) { delete s; return rc; }
    *out = reinterpret_cast<gplum_system *>(s);
    return GPLUM_OK;
}

int gplum_system_destroy(gplum_system *s) {
    delete reinterpret_cast<System *>(s);
    return GPLUM_OK;
}

int gplum_system_set_params(gplum_system *s, const gplum_params *p) {
    if (!s || !p) return gplum::fail(GPLUM_E_ARG, "null argument");
    reinterpret_cast<System *>(s)->p = *p;
    return GPLUM_OK;
}

int gplum_system_set_particles(gplum_system *s, const gplum_particle *pp, int n, double time_sys) {
    if (!s || (n > 0 && !pp) || n < 0) return gplum::fail(GPLUM_E_ARG, "bad particle array");
    return reinterpret_cast<System *>(s)->set_particles(pp, n, time_sys);
}

int gplum_system_get_particles(gplum_system *s, gplum_particle *pp, int n) {
    if (!s || !pp) return gplum::fail(GPLUM_E_ARG, "null argument");
    return reinterpret_cast<System *>(s)->get_particles(pp, n);
}

int gplum_system_count(gplum_system *s, int *n) {
    if (!s || !n) return gplum::fail(GPLUM_E_ARG, "null argument");
    *n = reinterpret_cast<System *>(s)->n;
    return GPLUM_OK;
}

int gplum_system_initialize(gplum_system *s) {
    if (!s) return gplum::fail(GPLUM_E_ARG, "null system");
    return reinterpret_cast<System *>(s)->initialize();
}

int gplum_system_soft_force(gplum_system *s, int initial, int use_f64) {
    if (!s) return gplum::fail(GPLUM_E_ARG, "null system");
    return reinterpret_cast<System *>(s)->soft_force(initial != 0, use_f64 != 0);
}

int gplum_system_set_radii(gplum_system *s) {
    if (!s) return gplum::fail(GPLUM_E_ARG, "null system");
    System *S = reinterpret_cast<System *>(s);
    int rc = S->set_radii();
    if (rc) return rc;
    GPLUM_HIP(hipStreamSynchronize(S->st));
    return gplum::radii_check(S->rad);
}

int gplum_system_step(gplum_system *s, int nsteps) {
    if (!s || nsteps < 0) return gplum::fail(GPLUM_E_ARG, "bad arguments");
    System *S = reinterpret_cast<System *>(s);
    for (int k = 0; k < nsteps; k++) {
        int rc = S->step(true);
        if (rc) return rc;
    }
    // the steps' last kernels (the energy sums) are complete when the call returns
    GPLUM_HIP(hipStreamSynchronize(S->st));
    return GPLUM_OK;
}

int gplum_system_energy(gplum_system *s, double out[6]) {
    if (!s || !out) return gplum::fail(GPLUM_E_ARG, "null argument");
    System *S = reinterpret_cast<System *>(s);
    int rc = S->energy(out);
    if (rc) return rc;
    GPLUM_HIP(hipStreamSynchronize(S->st));
    if ((rc = S->merge_resolve())) return rc;
    out[5] = S->edisp;
    return GPLUM_OK;
}

int gplum_system_neighbors(gplum_system *s, int32_t *offset, int32_t *index, int cap, int *total) {
    if (!s || !offset || !total) return gplum::fail(GPLUM_E_ARG, "null argument");
    System *S = reinterpret_cast<System *>(s);
    if (!S->have_nbr) return gplum::fail(GPLUM_E_STATE, "no neighbour graph for the current order");
    GPLUM_HIP(hipStreamSynchronize(S->st));
    int rc = gplum::nb_resolve(S->nb);
    if (rc) return rc;
    *total = S->nb.n_nbr;
    GPLUM_HIP(hipMemcpy(offset, S->nb.nbr_off.p, sizeof(int32_t) * (S->n + 1), hipMemcpyDeviceToHost));
    if (index && cap >= S->nb.n_nbr && S->nb.n_nbr > 0)
        GPLUM_HIP(hipMemcpy(index, S->nb.nbr_idx.p, sizeof(int32_t) * S->nb.n_nbr, hipMemcpyDeviceToHost));
    return GPLUM_OK;
}

int gplum_system_hard_part(gplum_system *s, int persistent) {
    if (!s) return gplum::fail(GPLUM_E_ARG, "null system");
    System *S = reinterpret_cast<System *>(s);
    int rc = S->hard_part(persistent != 0);
    if (!rc) { S->n_col_tot += S->step_ncol; S->edisp += S->step_edisp; }
    return rc;
}

int gplum_system_stats(gplum_system *s, gplum_stats *out) {
    if (!s || !out) return gplum::fail(GPLUM_E_ARG, "null argument");
    System *S = reinterpret_cast<System *>(s);
    int rc = S->fc_sync();
    if (rc) return rc;
    GPLUM_HIP(hipStreamSynchronize(S->st));
    S->resolve_timing();
    if ((rc = gplum::nb_resolve(S->nb))) return rc;
    if ((rc = S->merge_resolve())) return rc;
    if (S->fc.trained_rounds > 0 && (rc = gplum::fc_read_loss(S->fc, S->st_fc))) return rc;
    S->stats.fc_loss = S->fc.last_loss;
    *out = S->stats;
    unsigned long long pc[3] = {0ull, 0ull, 0ull};
    if (S->pair_cnt.p) {
        GPLUM_HIP(hipStreamSynchronize(S->st));
        GPLUM_HIP(hipMemcpy(pc, S->pair_cnt.p, sizeof(pc), hipMemcpyDeviceToHost));
    }
    out->n_epep_near = (int64_t)pc[0];
    out->n_epep_far = (int64_t)pc[1];
    out->n_epsp = (int64_t)pc[2];
    out->n_sp_entries = S->tree.n_sp_total;
    out->n_groups = S->last_groups;
    out->n_candidates = S->last_ncand;
    out->n_neighbors = S->nb.n_nbr;
    out->n_collisions = S->n_col_tot;
    out->n_removed = S->n_remove_tot;
    out->n_recomputes = S->n_recompute;
    out->time_sys = S->time_sys;
    return GPLUM_OK;
}

int gplum_system_set_boundary(gplum_system *s, double r_max, double r_min, int reset_step) {
    if (!s) return gplum::fail(GPLUM_E_ARG, "null system");
    if (!(r_max >= r_min) || !(r_min >= 0.)) return gplum::fail(GPLUM_E_ARG, "need 0 <= r_min <= r_max");
    if (reset_step < 0) return gplum::fail(GPLUM_E_ARG, "reset_step must be >= 0");
    System *S = reinterpret_cast<System *>(s);
    S->r_max = r_max;
    S->r_min = r_min;
    S->reset_step = reset_step;
    return GPLUM_OK;
}

int gplum_system_set_event_logs(gplum_system *s, const char *collision_path, const char *remove_path) {
    if (!s) return gplum::fail(GPLUM_E_ARG, "null system");
    System *S = reinterpret_cast<System *>(s);
    FILE *fc = nullptr, *fr = nullptr;
    if (collision_path && *collision_path && !(fc = std::fopen(collision_path, "a")))
        return gplum::fail(GPLUM_E_ARG, std::string("cannot open ") + collision_path);
    if (remove_path && *remove_path && !(fr = std::fopen(remove_path, "a"))) {
        if (fc) std::fclose(fc);
        return gplum::fail(GPLUM_E_ARG, std::string("cannot open ") + remove_path);
    }
    if (S->log_col) std::fclose(S->log_col);
    if (S->log_rem) std::fclose(S->log_rem);
    S->log_col = fc;
    S->log_rem = fr;
    return GPLUM_OK;
}

int gplum_system_reset_stats(gplum_system *s, int timing) {
    if (!s) return gplum::fail(GPLUM_E_ARG, "null system");
    System *S = reinterpret_cast<System *>(s);
    int rc = S->fc_sync();
    if (rc) return rc;
    GPLUM_HIP(hipStreamSynchronize(S->st));
    S->resolve_timing();
    std::memset(&S->stats, 0, sizeof(S->stats));
    S->timing = timing != 0;
    return GPLUM_OK;
}

int gplum_system_autotune(gplum_system *s, int trials, gplum_tuning *out) {
    if (!s) return gplum::fail(GPLUM_E_ARG, "null system");
    return reinterpret_cast<System *>(s)->autotune(trials, out);
}

int gplum_system_get_tuning(gplum_system *s, gplum_tuning *out) {
    if (!s || !out) return gplum::fail(GPLUM_E_ARG, "null argument");
    System *S = reinterpret_cast<System *>(s);
    *out = gplum_tuning{S->p.n_group_limit, S->soft_tile, S->hard_wg_per_cu, S->hard_lds_max, S->hard_giant_min,
                        S->tree.coop_min, S->hard_terms_max, 0, 0., 0.};
    return GPLUM_OK;
}

int gplum_system_set_tuning(gplum_system *s, const gplum_tuning *t) {
    if (!s || !t) return gplum::fail(GPLUM_E_ARG, "null argument");
    if (t->n_group_limit < 1 || t->n_group_limit > 256) return gplum::fail(GPLUM_E_ARG, "n_group_limit must be 1..256");
    if (t->soft_tile < 64 || t->soft_tile > 256 || t->soft_tile % 64)
        return gplum::fail(GPLUM_E_ARG, "soft_tile must be 64, 128, 192 or 256");
    if (t->hard_wg_per_cu < 1 || t->hard_wg_per_cu > 64) return gplum::fail(GPLUM_E_ARG, "hard_wg_per_cu must be 1..64");
    if (t->hard_lds_max < 64 || t->hard_lds_max > 256) return gplum::fail(GPLUM_E_ARG, "hard_lds_max must be 64..256");
    if (t->hard_giant_min < 64) return gplum::fail(GPLUM_E_ARG, "hard_giant_min must be >= 64");
    if (t->tree_coop_min < 0) return gplum::fail(GPLUM_E_ARG, "tree_coop_min must be >= 0");
    if (t->hard_terms_max < 0 || t->hard_terms_max > gplum::TERMS_MAX)
        return gplum::fail(GPLUM_E_ARG, "hard_terms_max must be 0.." + std::to_string(gplum::TERMS_MAX));
    System *S = reinterpret_cast<System *>(s);
    S->tree.coop_min = t->tree_coop_min > 0 ? t->tree_coop_min : (1 << 16);
    S->p.n_group_limit = t->n_group_limit;
    S->soft_tile = t->soft_tile;
    S->hard_wg_per_cu = t->hard_wg_per_cu;
    S->hard_lds_max = t->hard_lds_max;
    S->hard_giant_min = t->hard_giant_min;
    S->hard_terms_max = t->hard_terms_max;
    return GPLUM_OK;
}

int gplum_system_set_variant(gplum_system *s, int fp64_soft, int per_cluster_hard) {
    if (!s) return gplum::fail(GPLUM_E_ARG, "null system");
    System *S = reinterpret_cast<System *>(s);
    S->plain_f64 = fp64_soft != 0;
    S->plain_per_cluster = per_cluster_hard != 0;
    return GPLUM_OK;
}

int gplum_system_set_forecast(gplum_system *s, int mode, int train_steps, int batch, float lr) {
    if (!s || mode < 0 || mode > 2 || train_steps < 0 || batch <= 0 || !(lr > 0.f))
        return gplum::fail(GPLUM_E_ARG, "bad arguments");
    System *S = reinterpret_cast<System *>(s);
    int rc = S->fc_sync();
    if (rc) return rc;
    S->fc.enabled = mode != 0;
    S->fc.uniform = mode == 2;
    S->fc.train_steps = train_steps;
    S->fc.batch = batch;
    S->fc.net.lr = lr;
    return GPLUM_OK;
}

int gplum_system_set_forecast_every(gplum_system *s, int every) {
    if (!s || every < 1) return gplum::fail(GPLUM_E_ARG, "bad arguments");
    System *S = reinterpret_cast<System *>(s);
    int rc = S->fc_sync();
    if (rc) return rc;
    S->fc.predict_every = every;
    return GPLUM_OK;
}

int gplum_system_get_forecast(gplum_system *s, float *theta, float *mu, float *inv_sd, int *trained) {
    if (!s) return gplum::fail(GPLUM_E_ARG, "null system");
    System *S = reinterpret_cast<System *>(s);
    int rc = S->fc_sync();
    if (rc) return rc;
    if (trained) *trained = S->fc.trained_rounds;
    if (theta) {
        if (!S->fc.net.initialized) return gplum::fail(GPLUM_E_STATE, "forecaster has not been trained");
        GPLUM_HIP(hipMemcpy(theta, S->fc.net.theta.p, sizeof(float) * gplum::MLP_NPARAM, hipMemcpyDeviceToHost));
    }
    for (int q = 0; q < gplum::MLP_NIN; q++) {
        if (mu) mu[q] = S->fc.norm.mu[q];
        if (inv_sd) inv_sd[q] = S->fc.norm.inv_sd[q];
    }
    return GPLUM_OK;
}

int gplum_system_put_forecast(gplum_system *s, const float *theta, const float *mu, const float *inv_sd) {
    if (!s || !theta || !mu || !inv_sd) return gplum::fail(GPLUM_E_ARG, "null argument");
    System *S = reinterpret_cast<System *>(s);
    int rc = S->fc_sync();
    if (rc) return rc;
    if ((rc = S->fc.net.ensure(1))) return rc;
    GPLUM_HIP(hipMemcpy(S->fc.net.theta.p, theta, sizeof(float) * gplum::MLP_NPARAM, hipMemcpyHostToDevice));
    GPLUM_HIP(hipMemset(S->fc.net.m.p, 0, sizeof(float) * gplum::MLP_NPARAM));
    GPLUM_HIP(hipMemset(S->fc.net.v.p, 0, sizeof(float) * gplum::MLP_NPARAM));
    S->fc.net.adam_t = 0;
    S->fc.net.initialized = true;
    for (int q = 0; q < gplum::MLP_NIN; q++) {
        S->fc.norm.mu[q] = mu[q];
        S->fc.norm.inv_sd[q] = inv_sd[q];
    }
    S->fc.normalized = true;
    S->fc.trained_rounds = std::max(S->fc.trained_rounds, S->fc.min_rounds);
    return GPLUM_OK;
}

}  // extern "C"
