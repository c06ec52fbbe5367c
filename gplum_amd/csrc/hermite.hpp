// hermite.hpp -- batched FP64 Hermite (hard part) and isolated-particle drift.
#pragma once
#include "engine.hpp"

namespace gplum {

constexpr int HERMITE_BLOCK = 64;      // threads per workgroup, clusters up to HARD_SEQ_MAX
// threads per workgroup (8 waves, two per SIMD) of the larger clusters' kernels
constexpr int HERMITE_BIG_BLOCK = 512;
constexpr int HERMITE_LDS = 256;     // clusters up to this size keep j-data in LDS
// Clusters up to HARD_SEQ_MAX particles sum each particle's hard force sequentially in
// j (one lane per particle); larger ones with 64 lane-strided partial sums and a
// butterfly (all lanes on one particle), an order the oracle reproduces.
constexpr int HARD_SEQ_MAX = 64;
// Clusters of 2 .. TERMS_MAX bodies spread each block step over one lane per particle,
// term (Sun or partner) and component (4 n^2 <= 64 lanes).
constexpr int TERMS_MAX = 4;
// Doubles per collision record (perfect_merge, for the collision log).
constexpr int COLREC = 20;

// Gathered cluster particles (SoA, vectors interleaved x,y,z).
struct HardBuffers {
    DArray<double> pos, vel, acc_s, jerk_s, acc_d, jerk_d, xp, vp, a0_s, j0_s, a0_d, j0_d;
    DArray<double> a2_s, a3_s, a2_d, a3_d;
    DArray<double> mass, rinv, time, dt, acc0, r_planet, f, phi_d, phi_s;
    DArray<int64_t> id;
    DArray<int32_t> flags, twin_of, act;
    DArray<int> next, ncol;
    DArray<double> edisp;
    DArray<long long> nblock, npair;    // block steps, hard_pair evaluations per cluster
    DArray<unsigned char> g_state;      // giant-cluster loop state
    DArray<double> g_r;                 // giant-cluster collision partial minima
    DArray<int> g_i, g_j;
    // collision records of the last integration (crec_cap > 0: recorded; the caller
    // sizes the buffers and zeroes ccount before the launch)
    DArray<double> crec;
    DArray<unsigned> ccount;
    unsigned crec_cap = 0;
    int max_size = 0;
    int ensure(int n) {
        GPLUM_ENSURE(pos, 3 * n); GPLUM_ENSURE(vel, 3 * n); GPLUM_ENSURE(acc_s, 3 * n);
        GPLUM_ENSURE(jerk_s, 3 * n); GPLUM_ENSURE(acc_d, 3 * n); GPLUM_ENSURE(jerk_d, 3 * n);
        GPLUM_ENSURE(xp, 3 * n); GPLUM_ENSURE(vp, 3 * n); GPLUM_ENSURE(a0_s, 3 * n);
        GPLUM_ENSURE(j0_s, 3 * n); GPLUM_ENSURE(a0_d, 3 * n); GPLUM_ENSURE(j0_d, 3 * n);
        GPLUM_ENSURE(a2_s, 3 * n); GPLUM_ENSURE(a3_s, 3 * n); GPLUM_ENSURE(a2_d, 3 * n);
        GPLUM_ENSURE(a3_d, 3 * n);
        GPLUM_ENSURE(mass, n); GPLUM_ENSURE(rinv, n); GPLUM_ENSURE(time, n); GPLUM_ENSURE(dt, n);
        GPLUM_ENSURE(acc0, n); GPLUM_ENSURE(r_planet, n); GPLUM_ENSURE(f, n); GPLUM_ENSURE(phi_d, n);
        GPLUM_ENSURE(phi_s, n); GPLUM_ENSURE(id, n); GPLUM_ENSURE(flags, n); GPLUM_ENSURE(twin_of, n);
        GPLUM_ENSURE(act, n);
        return GPLUM_OK;
    }
};

struct IsoArgs {
    double *pos, *vel, *acc_d, *jerk_d, *acc_s, *jerk_s, *phi_d, *phi_s, *time, *dt, *acc0;
    const int32_t *number;
    int32_t *n_cluster;
};

// Integrates the ncl clusters whose particles are gathered in hb at offsets
// d_off[0..ncl] (local time 0 -> dt_tree; caller sets time = 0).  d_order (optional)
// gives the processing order (largest first).  persistent: one grid pulling
// clusters from an atomic queue; otherwise one workgroup per cluster.
// st2/ev_fork/ev_join (optional): clusters above HARD_SEQ_MAX run on st2 concurrently
// with the small-cluster kernel on st, and st waits for st2.  wg_per_cu: persistent
// workgroups per CU; lds_max (<= HERMITE_LDS): the largest cluster whose j-data is
// staged in LDS (larger ones read it from global memory); clusters above giant_min run
// over the whole GPU, several at a time on a stream pool, host-driven (not capturable in
// a graph).  terms_max: clusters of 2 .. terms_max bodies take the lane-per-term path
// (0: every cluster up to HARD_SEQ_MAX takes the one-lane-per-particle path).  None of
// these changes results.
constexpr int GIANT_MIN_DEFAULT = 1024;
constexpr int GIANT_OFF = 0x7fffffff;
int hard_launch(const gplum_params &p, HardBuffers &hb, int ncl, const int32_t *d_off,
                const int32_t *d_order, bool persistent, int n_cu, hipStream_t st, hipStream_t st2 = nullptr,
                hipEvent_t ev_fork = nullptr, hipEvent_t ev_join = nullptr, int wg_per_cu = 8,
                int lds_max = HERMITE_LDS, int giant_min = GIANT_MIN_DEFAULT, int terms_max = TERMS_MAX);
// Sun + hard acc/jerk and the initial block step of every cluster particle (as the
// initial force correction leaves them), for callers that bring bare clusters.
int hard_prepare(const gplum_params &p, HardBuffers &hb, int ncl, const int32_t *d_off, hipStream_t st);
// One launch for cluster c (n_c particles) only (the per-cluster-launch comparator).
int hard_launch_single(const gplum_params &p, HardBuffers &hb, int c, int n_c, const int32_t *d_off,
                       hipStream_t st);
// Isolated particles (number == 0): Kepler drift or Sun-only Hermite.
int isolated_launch(const gplum_params &p, const IsoArgs &a, int n, double t0, double t1,
                    hipStream_t st);

// dt_block's fast criterion against the correctly rounded one on n random states.
int hermite_dt_selfcheck(const gplum_params &p, long long n, uint64_t seed, long long *mismatches,
                         long long *fallbacks);

}  // namespace gplum
