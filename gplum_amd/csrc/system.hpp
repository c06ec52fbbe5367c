// system.hpp -- the device-resident P3T system (see system.hip).
#pragma once
#include "forecast.hpp"
#include "hermite.hpp"
#include "neighbor.hpp"

namespace gplum {

struct PartPtrs {
    double *pos, *vel, *acc, *acc_d, *jerk_d, *acc_s, *jerk_s, *mass, *r_out, *r_out_inv, *r_search;
    double *time, *dt, *acc0, *r_planet, *f, *phi, *phi_d, *phi_s;
    int64_t *id;
    int32_t *flags, *n_cluster, *number;
};

struct HardPtrs {
    double *pos, *vel, *acc_s, *jerk_s, *acc_d, *jerk_d, *mass, *rinv, *time, *dt, *acc0, *r_planet, *f;
    double *phi_d, *phi_s;
    int64_t *id;
    int32_t *flags;
};

enum {
    EV_TREE0, EV_TREE1, EV_FORCE0, EV_FORCE1, EV_NB0, EV_NB1, EV_HARD0, EV_HARD1,
    EV_FCA0, EV_FCA1, EV_FCB0, EV_FCB1, EV_FCC0, EV_FCC1, EV_HK0, EV_HK1, EV_COUNT
};

struct RadiiScratch {                // setCutoffRadii on the device (radii.hip)
    DArray<double> part, range, val, val_sorted, vd, csum;
    DArray<uint64_t> key, key_sorted;
    DArray<int> np, err, brange, cbase;
    DArray<unsigned char> tmp;
    HostBuf err_host;                 // the last call's error flag, read without a wait
    bool err_pending = false;
};
int device_cutoff_radii(const gplum_params &p, ParticleSoA &A, int n, RadiiScratch &R, hipStream_t st);
// the error flag of the last device_cutoff_radii, once the stream has passed it
int radii_check(RadiiScratch &R);

struct MergeScratch {                 // MergeParticle + boundary removal on the device (system.hip)
    DArray<int32_t> flag, scan, sel, keep, kscan, perm, rem, rscan, ridx, count;
    DArray<int32_t> tgt, twin_off, twin;
    DArray<double> ed, rdata;
    DArray<int64_t> rid;
    DArray<int64_t> sel_info;         // flags (as int32) and ids of the listed particles
    HostBuf ed_host;                  // the merged targets' energies, read back without a wait
    int ed_pending = 0;
};

// A particle that left the system through removeParticlesOutOfBoundary (func.h:209-380),
// as the remove log and the energy bookkeeping need it.
struct Removed {
    int64_t id;
    double time, mass, pos[3], vel[3], phi, phi_d, phi_s;
};

struct System {
    gplum_params p;
    hipStream_t st = nullptr, st2 = nullptr;   // st2: large clusters of the hard part
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // the forecaster's training runs on st_fc, overlapping the next step's hard part;
    // the next soft force (which rewrites its inputs and reads the network) waits for it
    hipStream_t st_fc = nullptr;
    hipStream_t st_iso = nullptr;       // the isolated particles' drift, beside the clusters
    hipEvent_t ev_iso_fork = nullptr, ev_iso_join = nullptr;
    hipEvent_t ev_fc_ready = nullptr, ev_fc_done = nullptr;
    bool fc_pending = false;
    hipEvent_t ev[EV_COUNT] = {};
    // Timed intervals outside the autotuner are not waited for: each tick records a fresh
    // event (evd), each interval waits in tpend until a synchronisation the step makes
    // anyway (or the statistics are read), so timing adds no host round trips.
    struct TimedPair {
        hipEvent_t a, b;
        double *dst;
    };
    hipEvent_t evd[EV_COUNT] = {};
    std::vector<hipEvent_t> ev_free;
    std::vector<TimedPair> tpend;
    int n_cu = 256;
    int n = 0;
    double time_sys = 0., edisp = 0., step_edisp = 0.;
    int n_col_tot = 0, step_ncol = 0;
    bool have_nbr = false, timing = false;
    bool plain_f64 = false, plain_per_cluster = false;   // comparison variant of step()
    // launch shapes picked by autotune() (defaults until it runs)
    int soft_tile = SOFT_TILE;          // j-particles staged per LDS pass of the soft kernel
    int hard_wg_per_cu = 8;             // persistent Hermite workgroups per CU
    int hard_lds_max = HERMITE_LDS;     // largest cluster whose j-data is staged in LDS
    int hard_giant_min = GIANT_MIN_DEFAULT;   // larger clusters run over the whole GPU
    int hard_terms_max = TERMS_MAX;     // clusters of 2 .. this many bodies: lane per term
    bool tuning = false;                // inside autotune(): no training, no statistics
    bool skip_train = false;            // the force recomputed after mergers: no training
    ParticleSoA A, B;
    RadiiScratch rad;
    MergeScratch mrg;
    Tree tree;
    DArray<JRec> jrec;
    DArray<SoftGroup> groups;
    DArray<int2> cand;
    DArray<unsigned> cand_count;
    unsigned cand_cap = 0, last_ncand = 0;
    CandSizer fc;                     // learned per-group candidate sizing
    NbState nb;
    ClusterList cl;
    Scratch scratch;
    HardBuffers hb;
    DArray<long long> hstat;          // k_hard_stats output
    DArray<int32_t> hstat_col;        // clusters with a collision (k_hard_stats)
    HostBuf hst;                      // pinned host side of the step's small reads
    DArray<double> time_c, eparts, eparts_step;
    double e_init[5] = {0., 0., 0., 0., 0.};
    // main loop (main_p3t.cpp:423-699): boundary of removeParticlesOutOfBoundary, the
    // periodic recompute interval and the step counter (read.h:111-115 defaults)
    double r_max = 40., r_min = 0.1;
    int reset_step = 1024;
    long long istep = 0;
    int n_remove_tot = 0, n_recompute = 0;
    FILE *log_rem = nullptr, *log_col = nullptr;   // remove / collision logs (append)
    long long last_groups = 0;
    DArray<unsigned long long> pair_cnt;   // near EP, far EP, SP pairs of the last force call
    gplum_stats stats{};

    explicit System(const gplum_params &prm);
    ~System();
    int init_device();
    int set_particles(const gplum_particle *pp, int count, double time);
    int get_particles(gplum_particle *pp, int count);
    int set_radii();
    int reorder();
    int gather_all(const int32_t *perm, int m, bool force_out = true);
    int soft_force(bool initial, bool use_f64);
    int fc_join();                      // st waits for the pending training (device side)
    int fc_sync();                      // the host waits for it
    int energy(double out[5]);
    int initialize();
    int kick();
    int hard_part(bool persistent);
    int merge_and_remove(int *n_removed);
    int energy_launch();
    int step(bool persistent);
    int hard_trial(int wg_per_cu, int lds_max, double *ms);
    int autotune(int trials, gplum_tuning *out);
    PartPtrs ptrs(ParticleSoA &s);
    HardPtrs hard_ptrs();
    int write_collision_log();
    void tick(int k);
    double elapsed(int a, int b);
    int merge_resolve();
    hipEvent_t timing_event();
    void accum(double *dst, int a, int b);
    void resolve_timing();
};

}  // namespace gplum
