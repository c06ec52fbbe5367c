/*
 * gplum_system.h -- device-resident P3T system: the particle state lives in HBM and
 * a whole soft step (main_p3t.cpp:423-699) runs as a chain of HIP kernels on one
 * stream.  The host exchanges particles in the gplum_particle layout below, which
 * carries the FPGrav fields the hot path reads and writes (src/particle.h:404-499).
 */
#ifndef GPLUM_SYSTEM_H
#define GPLUM_SYSTEM_H

#include <stdint.h>
#include "gplum_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GPLUM_FLAG_DEAD    1   /* FPGrav::isDead   */
#define GPLUM_FLAG_MERGED  2   /* FPGrav::isMerged */

/* One particle (FPGrav fields used by the P3T hot path). */
typedef struct gplum_particle {
    int64_t id;
    int64_t id_cluster;
    double  mass, r_planet, f;
    double  pos[3], vel[3];
    double  acc[3];                 /* soft acceleration, FPGrav::acc          */
    double  acc_d[3], jerk_d[3];    /* hard (neighbour) acceleration and jerk  */
    double  acc_s[3], jerk_s[3];    /* Sun acceleration and jerk               */
    double  phi, phi_d, phi_s;
    double  r_out, r_out_inv, r_search, v_disp;
    double  time, dt, acc0;
    gplum_neighbor_info neighbor;   /* FP32 candidates, then FP64 count        */
    int32_t n_cluster;
    int32_t flags;                  /* GPLUM_FLAG_*                            */
} gplum_particle;

/* Counters and timings of the system (timings accumulate while enabled by
 * gplum_system_reset_stats(s, 1); HIP events on the system's stream). */
typedef struct gplum_stats {
    int64_t steps;
    int64_t n_epep_near, n_epep_far, n_epsp;  /* pair evaluations of the last soft force */
    int64_t n_sp_entries;                     /* super-particles staged (sum over groups) */
    int64_t n_groups, n_candidates, n_neighbors;
    int64_t n_clusters, n_cluster_particles, max_cluster;
    int64_t block_steps;                      /* Hermite block steps, summed over clusters */
    int64_t n_collisions;
    int64_t cand_retries;                     /* candidate-buffer regrowths                */
    int64_t n_force_calls;
    /* learned candidate sizing (all summed over soft-force calls) */
    int64_t fc_calls;                         /* calls sized by the forecaster             */
    int64_t fc_slots;                         /* candidate slots planned                   */
    int64_t fc_ovf_groups;                    /* i-groups whose segment overflowed         */
    int64_t fc_ovf_candidates;                /* candidates redone by the fallback pass    */
    int64_t hard_pairs;                       /* Hermite pair evaluations (hard_pair calls) */
    int64_t max_block_steps;                  /* most block steps of one cluster (max over calls) */
    int64_t max_block_cluster;                /* members of that cluster                   */
    int64_t n_force_f64_calls;                /* soft forces run with the FP64 kernels     */
    int64_t n_removed;                        /* particles removed out of the boundary     */
    int64_t n_recomputes;                     /* steps ending with the Initial recompute   */
    double force_ms, tree_ms, finalize_ms, hard_ms;
    double time_sys;
    double fc_loss;                           /* last training loss (Poisson NLL)          */
    double fc_ms;                             /* features, forecast and fallback           */
    double hard_kernel_ms;                    /* the batched Hermite kernels (one launch set
                                                 per hard part), HIP events on their stream */
    int64_t n_hard_launches;                  /* launch sets timed in hard_kernel_ms       */
    double fc_train_ms;                       /* the forecaster's training (its own stream,
                                                 overlapping the next hard part)          */
} gplum_stats;

typedef struct gplum_system gplum_system;

/* Lifetime. */
int gplum_system_create(const gplum_params *p, gplum_system **out);
int gplum_system_destroy(gplum_system *s);
int gplum_system_set_params(gplum_system *s, const gplum_params *p);

/* Host <-> device particle exchange.  The system keeps particles in Morton order,
 * so get_particles returns them in the current device order (match by id). */
int gplum_system_set_particles(gplum_system *s, const gplum_particle *pp, int n, double time_sys);
int gplum_system_get_particles(gplum_system *s, gplum_particle *pp, int n);
int gplum_system_count(gplum_system *s, int *n);

/* Initialisation of main_p3t.cpp:289-364: time/flags/r_planet/f defaults,
 * setCutoffRadii (particle.h:1528-1543), then two rounds of
 * calcForceAllAndWriteBack + correctForceLongInitial (gravity_soft.h:374-528). */
int gplum_system_initialize(gplum_system *s);

/* setCutoffRadii alone (particle.h:1528-1543). */
int gplum_system_set_radii(gplum_system *s);

/* One tree force: replaces tree_grav.calcForceAllAndWriteBack(calcForceEPEPWithSearch,
 * calcForceEPSP, ...) (main_p3t.cpp:583-588) followed by correctForceLong
 * (gravity_soft.h:245-372), or correctForceLongInitial when initial != 0.
 * use_f64 runs the FP64 twin kernels for acc/phi (comparison runs). */
int gplum_system_soft_force(gplum_system *s, int initial, int use_f64);

/* Hard part of one step on the current neighbour graph: createNeighborCluster +
 * HardSystem::makeList + HardSystem::timeIntegrate (main_p3t.cpp:472-526;
 * hard.h:321-829).  persistent: one persistent batched launch (else one launch
 * per cluster). */
int gplum_system_hard_part(gplum_system *s, int persistent);

/* nsteps soft steps of the loop at main_p3t.cpp:423-699: kick, hard part, tree force +
 * correctForceLong, kick, calcEnergy, MergeParticle (func.h:160-207),
 * removeParticlesOutOfBoundary (func.h:209-380) and, after a collision, a removal or
 * every reset_step-th step (istep % reset_step == reset_step - 1, istep starting at
 * round(time_sys / dt_tree)), setCutoffRadii + tree force + correctForceLongInitial. */
int gplum_system_step(gplum_system *s, int nsteps);

/* The main loop's run parameters param.r_max, param.r_min and param.reset_step
 * (read.h:80-84, defaults 40, 0.1, 1024 at read.h:111-115): particles with |x| > r_max
 * or |x| < r_min leave the system at the end of a step, their energy booked as
 * dissipated (func.h:305-333); reset_step 0 disables the periodic recompute. */
int gplum_system_set_boundary(gplum_system *s, double r_max, double r_min, int reset_step);

/* The collision and remove logs (fout_col, fout_rem, main_p3t.cpp:379-395): files the
 * system appends a line to per merger (Collision0::write2File, collisionA.h:122-148) and
 * per removed particle (func.h:347-352), in the reference's formats.  NULL or "" closes
 * a log. */
int gplum_system_set_event_logs(gplum_system *s, const char *collision_path, const char *remove_path);

/* Energy::calcEnergy (energy.h:29-63): out = {etot, ekin, ephi_sun, ephi, ephi_d, edisp}. */
int gplum_system_energy(gplum_system *s, double out[6]);

/* Neighbour graph of the current order (NeighborList::n_list, neighbor.h:247-250):
 * offset[n+1]; index[offset[n]] when cap is large enough; *total = offset[n]. */
int gplum_system_neighbors(gplum_system *s, int32_t *offset, int32_t *index, int cap, int *total);

int gplum_system_stats(gplum_system *s, gplum_stats *out);
int gplum_system_reset_stats(gplum_system *s, int timing);

/* Comparison variant of gplum_system_step (no counterpart in the reference): fp64_soft
 * runs the soft force with the FP64 kernels (neighbour candidates still from the
 * FP32 predicate, so the neighbour graph is unchanged); per_cluster_hard integrates
 * the clusters with one launch per cluster instead of the persistent batched launch.
 * Both 0 (default) is the production path. */
int gplum_system_set_variant(gplum_system *s, int fp64_soft, int per_cluster_hard);

/* Launch shapes of the device system (no counterpart in the reference, whose i-group
 * size is the n_group_limit parameter of src/read.h:104).  gplum_system_autotune times,
 * on the system's current state and best of `trials` runs each, the soft stage (tree,
 * FP32 force, neighbour finalize) for n_group_limit in {64, 128, 256} x soft_tile in
 * {64, 128, 256}, and the Hermite kernels for hard_wg_per_cu in {2, 4, 8, 16} x
 * hard_lds_max in {128, 256}, keeps the fastest of each stage and recomputes the soft
 * force with it.  n_group_limit changes which tree nodes are opened (the tree
 * approximation, as the reference parameter does) and soft_tile the FP32 partial sums;
 * neighbour lists, clusters and every hard-part result are unchanged.  get/set read or apply a tuning (e.g. one cached from an earlier run);
 * soft_ms / hard_ms report the autotuner's measurements (0 from get). */
typedef struct gplum_tuning {
    int32_t n_group_limit;   /* i-group size of the tree walk                        */
    int32_t soft_tile;       /* j-particles per LDS pass of the FP32 soft kernel      */
    int32_t hard_wg_per_cu;  /* persistent Hermite workgroups per CU                  */
    int32_t hard_lds_max;    /* clusters up to this size stage their j-data in LDS    */
    int32_t hard_giant_min;  /* larger clusters run over the whole GPU (not tuned)    */
    int32_t tree_coop_min;   /* tree nodes of at least this many particles split with a
                                workgroup-wide search (not tuned; 0 = 65536)          */
    int32_t hard_terms_max;  /* clusters of 2 .. this many (<= 4) bodies spread a block
                                step over one lane per particle, term and component;
                                0: one lane per particle (not tuned; same results)   */
    int32_t pad_;
    double soft_ms, hard_ms; /* best measured soft stage / Hermite kernels, ms        */
} gplum_tuning;
int gplum_system_autotune(gplum_system *s, int trials, gplum_tuning *out);
int gplum_system_get_tuning(gplum_system *s, gplum_tuning *out);
int gplum_system_set_tuning(gplum_system *s, const gplum_tuning *t);

/* Learned candidate-buffer sizing (no counterpart in the reference, whose CPU lists
 * grow as std::vectors; see include/gplum_forecast.h).  mode 1 trains the forecaster
 * network (train_steps Adam steps of batch samples, learning rate lr, on its own stream
 * while the next hard part runs) and, once trained, sizes the EP-EP candidate buffer
 * per i-group from its forecast (on the matrix cores; the rates travel with the
 * particles through every reorder); a fresh
 * forecast and training round are made every `every` soft forces (default 4, see
 * gplum_system_set_forecast_every); overflowed groups are redone exactly, so particles, neighbour
 * lists and clusters are identical in every mode.  mode 2 is the same per-group
 * scheme with the network replaced by the previous call's mean count per particle
 * (the ablation the network is measured against); mode 0 (default) uses one global
 * buffer sized by the previous call's total and redoes the force on overflow. */
int gplum_system_set_forecast(gplum_system *s, int mode, int train_steps, int batch, float lr);
int gplum_system_set_forecast_every(gplum_system *s, int every);
/* Copy the network out (theta[4545] and the feature standardisation mu[4],
 * inv_sd[4]) or into another system, e.g. one started from a new snapshot. */
int gplum_system_get_forecast(gplum_system *s, float *theta, float *mu, float *inv_sd, int *trained);
int gplum_system_put_forecast(gplum_system *s, const float *theta, const float *mu, const float *inv_sd);

/* Hard integration of caller-provided clusters (host arrays): cluster c is
 * pp[offsets[c] .. offsets[c+1]).  Each cluster advances one soft step dt_tree as
 * timeIntegrate_multi does (src/hermite.h:224-479) from the particles' current
 * pos/vel/dt/acc0/r_out_inv (FPHard(FPGrav) local time, src/particle.h:1096-1116),
 * replacing the cluster loops of HardSystem::timeIntegrate (src/hard.h:583-788).
 * mode bit 0 set: one persistent batched launch (largest cluster first); clear: one
 * launch per cluster (plain comparator).  mode bit 2 (with bit 0): every cluster above
 * 64 particles runs on the whole-GPU path meant for giant clusters (normally above 1024;
 * same results).  mode bit 3 (with bit 0): clusters of 2 to 4 bodies run one lane per
 * particle instead of one lane per particle, term and component (same results).
 * mode bit 1 set: the clusters are bare
 * (pos, vel, mass, r_out_inv, acc0, dt only): Sun and hard acc/jerk and the first
 * block step are computed on the device first, as the initial force correction
 * would leave them (not timed).  Outputs (each may be null): edisp (collision
 * energy), block_steps (summed), kernel_ms (HIP events around the integration
 * launches), pair_evals (Hermite pair-force evaluations, summed). */
int gplum_hard_integrate(const gplum_params *p, gplum_particle *pp, int n, const int32_t *offsets,
                         int n_cluster, int mode, double *edisp, int64_t *block_steps,
                         double *kernel_ms, int64_t *pair_evals);

/* Snapshots in the reference's ASCII format: FileHeader::readAscii / writeAscii
 * (src/energy.h:89-116) and FPGrav::readAscii / writeAscii (src/particle.h:829-855),
 * same printf/scanf formats.  Energies: etot, ekin, ephi_sun, ephi_planet, edisp.
 * gplum_snapshot_read with pp == NULL reads the header only (n = n_body); fields not in
 * the file are zero, id_cluster = id.  gplum_system_snapshot writes the system's
 * particles with its initial and current energies (makeSnap, src/func.h:53-66).
 * gplum_system_energy_line formats the energy-log line of outputStep (src/func.h:87-130,
 * without the optional columns): time, n, etot, relative error, largest cluster, number
 * of clusters, isolated particles. */
typedef struct gplum_snapshot_header {
    double time;
    int32_t n_body, id_next;
    double e_init[5], e_now[5];
} gplum_snapshot_header;
int gplum_snapshot_read(const char *path, gplum_snapshot_header *h, gplum_particle *pp, int cap, int *n);
int gplum_snapshot_write(const char *path, const gplum_snapshot_header *h, const gplum_particle *pp, int n);
int gplum_system_snapshot(gplum_system *s, const char *path, int32_t id_next);
int gplum_system_energy_line(gplum_system *s, char *buf, int cap);

/* The binary restart file snap_tmp.dat (makeSnapTmp, src/func.h:68-80; read back at
 * main_p3t.cpp:208, 273-290): FDPS writeParticleBinary(filename, header) stores the
 * header's FileHeader::writeBinary (an fwrite of the object, src/energy.h:119-126) and
 * then every particle's FPGrav::writeBinary (an fwrite of the object, src/particle.h:
 * 868-876), in array order.  The records below are those objects' bytes in the
 * reference's default build (src/Makefile:29-36: USE_INDIVIDUAL_CUTOFF, USE_QUAD,
 * COLLISION; without polar coordinates, the 6th-order Sun, binary merging and a constant
 * random velocity), on x86-64 with natural alignment. */
typedef struct gplum_energy {          /* Energy, src/energy.h:3-11 (56 B) */
    double etot, ekin, ephi_sun, ephi_planet, ephi, ephi_d, edisp;
} gplum_energy;

typedef struct gplum_file_header {     /* FileHeader, src/energy.h:70-79 (128 B) */
    int32_t n_body, id_next;
    double time;
    gplum_energy e_init, e_now;
} gplum_file_header;

typedef struct gplum_fpgrav {          /* FPGrav, src/particle.h:93-170, 404-495 (344 B) */
    int32_t id_local, myrank;          /* EPIGrav                                        */
    double pos[3], r_out, r_search;
    int64_t id;                        /* EPJGrav                                        */
    double mass, vel[3], acc_d[3];
    double acc[3], acc_s[3], jerk_d[3], jerk_s[3], acc_gd[3];   /* FPGrav               */
    double phi, phi_d, phi_s, v_disp, r_out_inv, time, dt, acc0, r_planet, f;
    gplum_neighbor_info neighbor;
    int64_t id_cluster;
    int32_t n_cluster;
    uint8_t inDomain, isSent, isDead, isMerged;
} gplum_fpgrav;

/* gplum_restart_write converts pp to FPGrav records (id_local = array index, myrank 0,
 * acc_gd 0, inDomain 1, isSent 0, isDead/isMerged from flags); gplum_restart_read does
 * the reverse (pp == NULL: header only, *n = n_body).  A run restarts as the reference
 * does (main_p3t.cpp:273-290): time_sys = round(time / dt_tree) dt_tree, then
 * gplum_system_set_particles + gplum_system_initialize.  gplum_system_restart_write
 * writes the system's particles with its initial and current energies (makeSnapTmp). */
int gplum_restart_write(const char *path, const gplum_file_header *h, const gplum_particle *pp, int n);
int gplum_restart_read(const char *path, gplum_file_header *h, gplum_particle *pp, int cap, int *n);
int gplum_system_restart_write(gplum_system *s, const char *path, int32_t id_next);

/* The persistent form of gplum_hard_integrate captured into a hipGraph and replayed
 * `replays` times (no counterpart in the reference; same clusters, results and
 * outputs as mode 1, for callers that replay a fixed batch).  The graph restores the
 * uploaded state, resets the cluster queue with a kernel and launches the size-class
 * kernels on two streams; every replay therefore integrates the same step.  prepare
 * as mode bit 1.  Outputs (each may be null): ms_per_replay (HIP events around each
 * replay, mean), mismatches (replays whose outputs differ bit for bit from the first
 * replay's: 0 expected).  pp receives the first replay's result. */
int gplum_hard_integrate_graph(const gplum_params *p, gplum_particle *pp, int n, const int32_t *offsets,
                               int n_cluster, int prepare, int replays, double *edisp, int64_t *block_steps,
                               double *ms_per_replay, int *mismatches);

/* Self-check of the Hermite time-step rule (no counterpart in the reference; for tests
 * and for users who port the library to other hardware).  FPHard::calcDeltat
 * (particle.h:1276-1314) compares the criterion dt_1 with the candidate steps only;
 * the kernels evaluate dt_1 with the hardware square root and reciprocal and fall back
 * to the correctly rounded evaluation when the fast value lies within 2^-10 of a
 * candidate.  This entry evaluates both forms on n random states (seeded) on the
 * device and reports the results that differ in any bit (0 expected) and the fast
 * values that took the fallback. */
int gplum_hermite_dt_selfcheck(const gplum_params *p, int64_t n, uint64_t seed, int64_t *mismatches,
                               int64_t *fallbacks);

#ifdef __cplusplus
}
#endif
#endif
