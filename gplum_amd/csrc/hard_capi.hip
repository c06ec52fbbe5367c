// hard_capi.hip -- batched hard integration on caller-provided clusters (host arrays).
//
// Replaces the per-cluster loop of HardSystem::timeIntegrate (src/hard.h:583-788):
// every cluster [off[c], off[c+1]) of pp is integrated over one soft step with
// timeIntegrate_multi semantics (src/hermite.h:224-479), in one persistent launch
// (mode 1) or with one launch per cluster (mode 0, the plain comparator).
#include <vector>

#include "hermite.hpp"
#include "system.hpp"

namespace gplum {
namespace {

struct HardCtx {
    hipStream_t st = nullptr, st2 = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, fork = nullptr, join = nullptr;
    HardBuffers hb;
    DArray<int32_t> off, order;
    int n_cu = 256;
    int init() {
        if (st) return GPLUM_OK;
        int count = 0, dev = 0;
        GPLUM_HIP(hipGetDeviceCount(&count));
        if (count <= 0) return fail(GPLUM_E_NODEVICE, "no HIP device");
        GPLUM_HIP(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        GPLUM_HIP(hipGetDeviceProperties(&prop, dev));
        n_cu = prop.multiProcessorCount;
        GPLUM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        GPLUM_HIP(hipStreamCreateWithFlags(&st2, hipStreamNonBlocking));
        GPLUM_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        GPLUM_HIP(hipEventCreateWithFlags(&join, hipEventDisableTiming));
        GPLUM_HIP(hipEventCreate(&e0));
        GPLUM_HIP(hipEventCreate(&e1));
        return GPLUM_OK;
    }
};

thread_local HardCtx t_hard;

template <class T>
int upload(DArray<T> &d, const std::vector<T> &h, hipStream_t st) {
    GPLUM_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, st));
    return GPLUM_OK;
}
template <class T>
int download(std::vector<T> &h, const DArray<T> &d, hipStream_t st) {
    GPLUM_HIP(hipMemcpyAsync(h.data(), d.p, sizeof(T) * h.size(), hipMemcpyDeviceToHost, st));
    return GPLUM_OK;
}

}  // namespace
}  // namespace gplum

using namespace gplum;

extern "C" int gplum_hard_integrate(const gplum_params *p, gplum_particle *pp, int n, const int32_t *offsets,
                                    int n_cluster, int mode, double *edisp, int64_t *block_steps,
                                    double *kernel_ms, int64_t *pair_evals) {
    if (!p || (n > 0 && (!pp || !offsets)) || n < 0 || n_cluster < 0)
        return fail(GPLUM_E_ARG, "bad arguments");
    if (n_cluster == 0) return GPLUM_OK;
    if (offsets[0] != 0 || offsets[n_cluster] != n) return fail(GPLUM_E_ARG, "offsets must span [0, n)");
    HardCtx &c = t_hard;
    int rc = c.init();
    if (rc) return rc;
    HardBuffers &hb = c.hb;
    if ((rc = hb.ensure(n))) return rc;
    int max_size = 0;
    std::vector<int32_t> order(n_cluster), sizes(n_cluster);
    for (int k = 0; k < n_cluster; k++) {
        sizes[k] = offsets[k + 1] - offsets[k];
        if (sizes[k] <= 0) return fail(GPLUM_E_ARG, "empty cluster");
        max_size = std::max(max_size, sizes[k]);
        order[k] = k;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sizes[a] > sizes[b]; });
    hb.max_size = max_size;
    std::vector<double> v3(3 * (size_t)n), v1(n), tc(n);
    std::vector<int64_t> ids(n);
    std::vector<int32_t> fl(n, 0), offs(offsets, offsets + n_cluster + 1);
    auto put3 = [&](DArray<double> &d, size_t fo) {
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) v3[3 * i + k] = *(const double *)((const char *)&pp[i] + fo + 8 * k);
        return upload(d, v3, c.st);
    };
    auto put1 = [&](DArray<double> &d, size_t fo) {
        for (int i = 0; i < n; i++) v1[i] = *(const double *)((const char *)&pp[i] + fo);
        return upload(d, v1, c.st);
    };
#define P3(dst, f) if ((rc = put3(hb.dst, offsetof(gplum_particle, f)))) return rc
#define P1(dst, f) if ((rc = put1(hb.dst, offsetof(gplum_particle, f)))) return rc
    P3(pos, pos); P3(vel, vel); P3(acc_s, acc_s); P3(jerk_s, jerk_s); P3(acc_d, acc_d); P3(jerk_d, jerk_d);
    P1(mass, mass); P1(rinv, r_out_inv); P1(dt, dt); P1(acc0, acc0); P1(r_planet, r_planet); P1(f, f);
#undef P3
#undef P1
    for (int i = 0; i < n; i++) { tc[i] = pp[i].time; v1[i] = 0.; ids[i] = pp[i].id; }
    if ((rc = upload(hb.time, v1, c.st))) return rc;
    if ((rc = upload(hb.id, ids, c.st))) return rc;
    if ((rc = upload(hb.flags, fl, c.st))) return rc;
    GPLUM_ENSURE(c.off, n_cluster + 1);
    GPLUM_ENSURE(c.order, n_cluster);
    if ((rc = upload(c.off, offs, c.st))) return rc;
    if ((rc = upload(c.order, order, c.st))) return rc;
    GPLUM_ENSURE(hb.next, 3);
    GPLUM_ENSURE(hb.edisp, n_cluster);
    GPLUM_ENSURE(hb.ncol, n_cluster);
    GPLUM_ENSURE(hb.nblock, n_cluster);
    GPLUM_ENSURE(hb.npair, n_cluster);
    if (mode & 2) {
        // bare clusters: time is the local time 0 here, dt the caller's previous step
        if ((rc = upload(hb.time, v1, c.st))) return rc;
        if ((rc = hard_prepare(*p, hb, n_cluster, c.off.p, c.st))) return rc;
    }
    GPLUM_HIP(hipEventRecord(c.e0, c.st));
    if (mode & 1) {
        if ((rc = hard_launch(*p, hb, n_cluster, c.off.p, c.order.p, true, c.n_cu, c.st, c.st2, c.fork, c.join)))
            return rc;
    } else {
        for (int k = 0; k < n_cluster; k++)
            if ((rc = hard_launch_single(*p, hb, k, sizes[k], c.off.p, c.st))) return rc;
    }
    GPLUM_HIP(hipEventRecord(c.e1, c.st));
    GPLUM_HIP(hipStreamSynchronize(c.st));
    float ms = 0.f;
    GPLUM_HIP(hipEventElapsedTime(&ms, c.e0, c.e1));
    if (kernel_ms) *kernel_ms = ms;
    std::vector<double> ed(n_cluster);
    std::vector<long long> nb(n_cluster);
    if ((rc = download(ed, hb.edisp, c.st))) return rc;
    if ((rc = download(nb, hb.nblock, c.st))) return rc;
    std::vector<long long> np(n_cluster);
    if ((rc = download(np, hb.npair, c.st))) return rc;
    auto get3 = [&](DArray<double> &d, size_t fo) {
        int r = download(v3, d, c.st);
        if (r) return r;
        GPLUM_HIP(hipStreamSynchronize(c.st));
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) *(double *)((char *)&pp[i] + fo + 8 * k) = v3[3 * i + k];
        return (int)GPLUM_OK;
    };
    auto get1 = [&](DArray<double> &d, size_t fo) {
        int r = download(v1, d, c.st);
        if (r) return r;
        GPLUM_HIP(hipStreamSynchronize(c.st));
        for (int i = 0; i < n; i++) *(double *)((char *)&pp[i] + fo) = v1[i];
        return (int)GPLUM_OK;
    };
#define G3(src, f) if ((rc = get3(hb.src, offsetof(gplum_particle, f)))) return rc
#define G1(src, f) if ((rc = get1(hb.src, offsetof(gplum_particle, f)))) return rc
    G3(pos, pos); G3(vel, vel); G3(acc_s, acc_s); G3(jerk_s, jerk_s); G3(acc_d, acc_d); G3(jerk_d, jerk_d);
    G1(dt, dt); G1(r_planet, r_planet); G1(phi_d, phi_d); G1(phi_s, phi_s); G1(time, time);
#undef G3
#undef G1
    if ((rc = download(ids, hb.id, c.st))) return rc;
    if ((rc = download(fl, hb.flags, c.st))) return rc;
    GPLUM_HIP(hipStreamSynchronize(c.st));
    double etot = 0.;
    long long btot = 0, ptot = 0;
    for (int k = 0; k < n_cluster; k++) {
        if (nb[k] < 0) return fail(GPLUM_E_STATE, "hard cluster exceeded the block-step cap");
        etot += ed[k];
        btot += nb[k];
        ptot += np[k];
    }
    for (int k = 0; k < n_cluster; k++)
        for (int i = offsets[k]; i < offsets[k + 1]; i++) {
            pp[i].time = tc[i] + pp[i].time;   // resetTime
            pp[i].id = ids[i];
            pp[i].flags = fl[i];
            pp[i].n_cluster = sizes[k];
        }
    if (edisp) *edisp = etot;
    if (block_steps) *block_steps = btot;
    if (pair_evals) *pair_evals = ptot;
    return GPLUM_OK;
}
