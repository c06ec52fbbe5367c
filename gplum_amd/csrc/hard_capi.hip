// hard_capi.hip -- batched hard integration on caller-provided clusters (host arrays).
//
// Replaces the per-cluster loop of HardSystem::timeIntegrate (src/hard.h:583-788):
// every cluster [off[c], off[c+1]) of pp is integrated over one soft step with
// timeIntegrate_multi semantics (src/hermite.h:224-479), in one persistent launch
// (mode 1) or with one launch per cluster (mode 0, the plain comparator).
// gplum_hard_integrate_graph captures the persistent launch into a hipGraph and
// replays it.
#include <cstring>
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
    // uploaded state kept for graph replays (each replay restores it first)
    DArray<double> b_pos, b_vel, b_acc_s, b_jerk_s, b_acc_d, b_jerk_d, b_mass, b_rinv, b_time, b_dt, b_acc0,
        b_r_planet, b_f;
    DArray<int64_t> b_id;
    DArray<int32_t> b_flags;
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


// Host-side layout of one call: cluster sizes, processing order (largest first) and the
// callers' times (the device integrates in local time from 0).
struct HardPlan {
    std::vector<int32_t> sizes, order;
    std::vector<double> tc;
    int n = 0, n_cluster = 0;
};

int check_args(const gplum_params *p, const gplum_particle *pp, int n, const int32_t *offsets, int n_cluster) {
    if (!p || (n > 0 && (!pp || !offsets)) || n < 0 || n_cluster < 0) return fail(GPLUM_E_ARG, "bad arguments");
    if (n_cluster > 0 && (offsets[0] != 0 || offsets[n_cluster] != n))
        return fail(GPLUM_E_ARG, "offsets must span [0, n)");
    return GPLUM_OK;
}

// Uploads the clusters into c.hb (and the offsets / order), then, for bare clusters,
// computes their acc/jerk and first block step on the device.
int hard_upload(HardCtx &c, const gplum_params *p, const gplum_particle *pp, int n, const int32_t *offsets,
                int n_cluster, bool prepare, HardPlan &pl) {
    int rc;
    HardBuffers &hb = c.hb;
    if ((rc = hb.ensure(n))) return rc;
    pl.n = n;
    pl.n_cluster = n_cluster;
    pl.sizes.assign(n_cluster, 0);
    pl.order.assign(n_cluster, 0);
    int max_size = 0;
    for (int k = 0; k < n_cluster; k++) {
        pl.sizes[k] = offsets[k + 1] - offsets[k];
        if (pl.sizes[k] <= 0) return fail(GPLUM_E_ARG, "empty cluster");
        max_size = std::max(max_size, pl.sizes[k]);
        pl.order[k] = k;
    }
    std::stable_sort(pl.order.begin(), pl.order.end(), [&](int a, int b) { return pl.sizes[a] > pl.sizes[b]; });
    hb.max_size = max_size;
    std::vector<double> v3(3 * (size_t)n), v1(n);
    std::vector<int64_t> ids(n);
    std::vector<int32_t> fl(n, 0), offs(offsets, offsets + n_cluster + 1);
    auto put3 = [&](DArray<double> &d, size_t fo) {
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) v3[3 * i + k] = *(const double *)((const char *)&pp[i] + fo + 8 * k);
        int r = upload(d, v3, c.st);
        if (r) return r;
        GPLUM_HIP(hipStreamSynchronize(c.st));   // v3 is reused
        return (int)GPLUM_OK;
    };
    auto put1 = [&](DArray<double> &d, size_t fo) {
        for (int i = 0; i < n; i++) v1[i] = *(const double *)((const char *)&pp[i] + fo);
        int r = upload(d, v1, c.st);
        if (r) return r;
        GPLUM_HIP(hipStreamSynchronize(c.st));   // v1 is reused
        return (int)GPLUM_OK;
    };
#define P3(dst, f) if ((rc = put3(hb.dst, offsetof(gplum_particle, f)))) return rc
#define P1(dst, f) if ((rc = put1(hb.dst, offsetof(gplum_particle, f)))) return rc
    P3(pos, pos); P3(vel, vel); P3(acc_s, acc_s); P3(jerk_s, jerk_s); P3(acc_d, acc_d); P3(jerk_d, jerk_d);
    P1(mass, mass); P1(rinv, r_out_inv); P1(dt, dt); P1(acc0, acc0); P1(r_planet, r_planet); P1(f, f);
#undef P3
#undef P1
    pl.tc.assign(n, 0.);
    for (int i = 0; i < n; i++) { pl.tc[i] = pp[i].time; v1[i] = 0.; ids[i] = pp[i].id; }
    if ((rc = upload(hb.time, v1, c.st))) return rc;
    if ((rc = upload(hb.id, ids, c.st))) return rc;
    if ((rc = upload(hb.flags, fl, c.st))) return rc;
    GPLUM_ENSURE(c.off, n_cluster + 1);
    GPLUM_ENSURE(c.order, n_cluster);
    if ((rc = upload(c.off, offs, c.st))) return rc;
    if ((rc = upload(c.order, pl.order, c.st))) return rc;
    GPLUM_ENSURE(hb.next, 3);
    GPLUM_ENSURE(hb.edisp, n_cluster);
    GPLUM_ENSURE(hb.ncol, n_cluster);
    GPLUM_ENSURE(hb.nblock, n_cluster);
    GPLUM_ENSURE(hb.npair, n_cluster);
    if (prepare && (rc = hard_prepare(*p, hb, n_cluster, c.off.p, c.st))) return rc;
    GPLUM_HIP(hipStreamSynchronize(c.st));
    return GPLUM_OK;
}

// Downloads the integrated clusters into pp and sums the per-cluster counters.
int hard_download(HardCtx &c, gplum_particle *pp, const int32_t *offsets, const HardPlan &pl, double *edisp,
                  int64_t *block_steps, int64_t *pair_evals) {
    int rc;
    const int n = pl.n, n_cluster = pl.n_cluster;
    HardBuffers &hb = c.hb;
    std::vector<double> v3(3 * (size_t)n), v1(n);
    std::vector<int64_t> ids(n);
    std::vector<int32_t> fl(n);
    std::vector<double> ed(n_cluster);
    std::vector<long long> nb(n_cluster), np(n_cluster);
    if ((rc = download(ed, hb.edisp, c.st))) return rc;
    if ((rc = download(nb, hb.nblock, c.st))) return rc;
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
    G1(dt, dt); G1(r_planet, r_planet); G1(phi_d, phi_d); G1(phi_s, phi_s); G1(time, time); G1(mass, mass);
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
            pp[i].time = pl.tc[i] + pp[i].time;   // resetTime
            pp[i].id = ids[i];
            pp[i].flags = fl[i];
            pp[i].n_cluster = pl.sizes[k];
        }
    if (edisp) *edisp = etot;
    if (block_steps) *block_steps = btot;
    if (pair_evals) *pair_evals = ptot;
    return GPLUM_OK;
}

// Every array the integration writes, as raw bytes (the replays of a graph must agree
// bit for bit).
int hard_outputs(HardCtx &c, const HardPlan &pl, std::vector<char> &out) {
    HardBuffers &hb = c.hb;
    const size_t n = pl.n, nc = pl.n_cluster;
    struct Seg { const void *p; size_t bytes; };
    const Seg segs[] = {{hb.pos.p, 24 * n}, {hb.vel.p, 24 * n}, {hb.acc_s.p, 24 * n}, {hb.jerk_s.p, 24 * n},
                        {hb.acc_d.p, 24 * n}, {hb.jerk_d.p, 24 * n}, {hb.dt.p, 8 * n}, {hb.time.p, 8 * n},
                        {hb.phi_d.p, 8 * n}, {hb.phi_s.p, 8 * n}, {hb.mass.p, 8 * n}, {hb.r_planet.p, 8 * n},
                        {hb.id.p, 8 * n}, {hb.flags.p, 4 * n}, {hb.edisp.p, 8 * nc}, {hb.nblock.p, 8 * nc}};
    size_t total = 0;
    for (const Seg &g : segs) total += g.bytes;
    out.resize(total);
    size_t at = 0;
    for (const Seg &g : segs) {
        GPLUM_HIP(hipMemcpyAsync(out.data() + at, g.p, g.bytes, hipMemcpyDeviceToHost, c.st));
        at += g.bytes;
    }
    GPLUM_HIP(hipStreamSynchronize(c.st));
    return GPLUM_OK;
}

}  // namespace
}  // namespace gplum

using namespace gplum;

extern "C" int gplum_hard_integrate(const gplum_params *p, gplum_particle *pp, int n, const int32_t *offsets,
                                    int n_cluster, int mode, double *edisp, int64_t *block_steps,
                                    double *kernel_ms, int64_t *pair_evals) {
    int rc = check_args(p, pp, n, offsets, n_cluster);
    if (rc) return rc;
    if (n_cluster == 0) return GPLUM_OK;
    HardCtx &c = t_hard;
    if ((rc = c.init())) return rc;
    HardPlan pl;
    if ((rc = hard_upload(c, p, pp, n, offsets, n_cluster, (mode & 2) != 0, pl))) return rc;
    GPLUM_HIP(hipEventRecord(c.e0, c.st));
    if (mode & 1) {
        // mode bit 2: every cluster above HARD_SEQ_MAX particles on the whole-GPU path;
        // mode bit 3: clusters of 2 to 4 bodies on the one-lane-per-particle path too
        const int giant_min = (mode & 4) ? HARD_SEQ_MAX : GIANT_MIN_DEFAULT;
        const int terms_max = (mode & 8) ? 0 : TERMS_MAX;
        if ((rc = hard_launch(*p, c.hb, n_cluster, c.off.p, c.order.p, true, c.n_cu, c.st, c.st2, c.fork, c.join, 8,
                              HERMITE_LDS, giant_min, terms_max)))
            return rc;
    } else {
        for (int k = 0; k < n_cluster; k++)
            if ((rc = hard_launch_single(*p, c.hb, k, pl.sizes[k], c.off.p, c.st))) return rc;
    }
    GPLUM_HIP(hipEventRecord(c.e1, c.st));
    GPLUM_HIP(hipStreamSynchronize(c.st));
    float ms = 0.f;
    GPLUM_HIP(hipEventElapsedTime(&ms, c.e0, c.e1));
    if (kernel_ms) *kernel_ms = ms;
    return hard_download(c, pp, offsets, pl, edisp, block_steps, pair_evals);
}

extern "C" int gplum_hard_integrate_graph(const gplum_params *p, gplum_particle *pp, int n, const int32_t *offsets,
                                          int n_cluster, int prepare, int replays, double *edisp,
                                          int64_t *block_steps, double *ms_per_replay, int *mismatches) {
    int rc = check_args(p, pp, n, offsets, n_cluster);
    if (rc) return rc;
    if (replays < 1) return fail(GPLUM_E_ARG, "replays must be >= 1");
    if (n_cluster == 0) return GPLUM_OK;
    HardCtx &c = t_hard;
    if ((rc = c.init())) return rc;
    HardPlan pl;
    if ((rc = hard_upload(c, p, pp, n, offsets, n_cluster, prepare != 0, pl))) return rc;
    HardBuffers &hb = c.hb;
    // keep the uploaded state: every replay starts from it
    struct Pair { void *dst, *src; size_t bytes; };
    GPLUM_ENSURE(c.b_pos, 3 * n); GPLUM_ENSURE(c.b_vel, 3 * n); GPLUM_ENSURE(c.b_acc_s, 3 * n);
    GPLUM_ENSURE(c.b_jerk_s, 3 * n); GPLUM_ENSURE(c.b_acc_d, 3 * n); GPLUM_ENSURE(c.b_jerk_d, 3 * n);
    GPLUM_ENSURE(c.b_mass, n); GPLUM_ENSURE(c.b_rinv, n); GPLUM_ENSURE(c.b_time, n); GPLUM_ENSURE(c.b_dt, n);
    GPLUM_ENSURE(c.b_acc0, n); GPLUM_ENSURE(c.b_r_planet, n); GPLUM_ENSURE(c.b_f, n); GPLUM_ENSURE(c.b_id, n);
    GPLUM_ENSURE(c.b_flags, n);
    const size_t N = (size_t)n;
    const Pair state[] = {{hb.pos.p, c.b_pos.p, 24 * N}, {hb.vel.p, c.b_vel.p, 24 * N},
                          {hb.acc_s.p, c.b_acc_s.p, 24 * N}, {hb.jerk_s.p, c.b_jerk_s.p, 24 * N},
                          {hb.acc_d.p, c.b_acc_d.p, 24 * N}, {hb.jerk_d.p, c.b_jerk_d.p, 24 * N},
                          {hb.mass.p, c.b_mass.p, 8 * N}, {hb.rinv.p, c.b_rinv.p, 8 * N},
                          {hb.time.p, c.b_time.p, 8 * N}, {hb.dt.p, c.b_dt.p, 8 * N},
                          {hb.acc0.p, c.b_acc0.p, 8 * N}, {hb.r_planet.p, c.b_r_planet.p, 8 * N},
                          {hb.f.p, c.b_f.p, 8 * N}, {hb.id.p, c.b_id.p, 8 * N}, {hb.flags.p, c.b_flags.p, 4 * N}};
    for (const Pair &s : state)
        GPLUM_HIP(hipMemcpyAsync(s.src, s.dst, s.bytes, hipMemcpyDeviceToDevice, c.st));
    GPLUM_HIP(hipStreamSynchronize(c.st));
    // capture: restore the state, reset the queue, the size-class kernels on two streams
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    GPLUM_HIP(hipStreamBeginCapture(c.st, hipStreamCaptureModeThreadLocal));
    for (const Pair &s : state)
        if (!rc && hipMemcpyAsync(s.dst, s.src, s.bytes, hipMemcpyDeviceToDevice, c.st) != hipSuccess)
            rc = fail(GPLUM_E_HIP, "hipMemcpyAsync failed during graph capture");
    // (the whole-GPU path for giant clusters is host-driven and cannot be captured)
    if (!rc)
        rc = hard_launch(*p, hb, n_cluster, c.off.p, c.order.p, true, c.n_cu, c.st, c.st2, c.fork, c.join, 8,
                         HERMITE_LDS, GIANT_OFF);
    const hipError_t ec = hipStreamEndCapture(c.st, &graph);
    if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
    GPLUM_HIP(ec);
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    GPLUM_HIP(ei);
    std::vector<char> first, again;
    double total_ms = 0.;
    int bad = 0;
    for (int r = 0; r < replays && !rc; r++) {
        GPLUM_HIP(hipEventRecord(c.e0, c.st));
        if (hipGraphLaunch(exec, c.st) != hipSuccess) { rc = fail(GPLUM_E_HIP, "hipGraphLaunch failed"); break; }
        GPLUM_HIP(hipEventRecord(c.e1, c.st));
        GPLUM_HIP(hipStreamSynchronize(c.st));
        float ms = 0.f;
        GPLUM_HIP(hipEventElapsedTime(&ms, c.e0, c.e1));
        total_ms += ms;
        if ((rc = hard_outputs(c, pl, r == 0 ? first : again))) break;
        if (r > 0 && (again.size() != first.size() || std::memcmp(again.data(), first.data(), first.size()) != 0))
            bad++;
    }
    (void)hipGraphExecDestroy(exec);
    if (rc) return rc;
    if (ms_per_replay) *ms_per_replay = total_ms / replays;
    if (mismatches) *mismatches = bad;
    return hard_download(c, pp, offsets, pl, edisp, block_steps, nullptr);
}

extern "C" int gplum_hermite_dt_selfcheck(const gplum_params *p, int64_t n, uint64_t seed, int64_t *mismatches,
                                          int64_t *fallbacks) {
    if (!p) return gplum::fail(GPLUM_E_ARG, "null params");
    long long m = 0, f = 0;
    const int rc = gplum::hermite_dt_selfcheck(*p, (long long)n, seed, &m, &f);
    if (mismatches) *mismatches = m;
    if (fallbacks) *fallbacks = f;
    return rc;
}
