/*
 * oracle.h -- CPU restatement of GPLUM's P3T hot path, used ONLY as the checker
 * (tests/, __graft_entry__.smoke()) and as the CPU baseline leg of bench.py.
 * It is test infrastructure: the product (libgplum_gpu.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Reference paths are
 * relative to the GPLUM source tree (src/...).
 *
 * Floating-point conventions (build with -ffp-contract=off):
 *  - FP64 code uses the reference's operation order with no fused multiply-add,
 *    as the reference's default build (no -mfma) does.
 *  - The FP32 soft kernels use a FIXED operation order (documented at
 *    orc_epep_f32) that the HIP kernels reproduce, so the FP32 neighbour predicate
 *    is bit-identical between the two.
 */
#ifndef GPLUM_ORACLE_H
#define GPLUM_ORACLE_H

#include <stdint.h>
#include "../include/gplum_gpu.h"
#include "../include/gplum_system.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reference parameter defaults (particle.h:917-961) and FPGrav::setGamma (:619-633). */
void orc_params_default(gplum_params *p);
void orc_params_set_gamma(gplum_params *p, double g);

/* ----------------------------------------------------------- src/cutfunc.h */
double orc_cutoff_f(const gplum_params *p, double y);
double orc_cutoff_W(const gplum_params *p, double rij, double r_out_inv);
double orc_cutoff_Kx(double x);
double orc_cutoff_K(const gplum_params *p, double rij, double r_out_inv);
double orc_cutoff_dKdr(const gplum_params *p, double x, double r_out_inv);
double orc_cutoff_dKdt(const gplum_params *p, double rij, double r_out_inv, double alpha);

/* ------------------------------------------------- soft kernels (PIKG order) */
void orc_force_clear(gplum_force *f, int n);
void orc_epep_f32(const gplum_epi *epi, int ni, const gplum_epj *epj, int nj,
                  gplum_force *force, float eps2);
void orc_epsp_quad_f32(const gplum_epi *epi, int ni, const gplum_spj_quad *spj, int nj,
                       gplum_force *force, float eps2);
void orc_epsp_mono_f32(const gplum_epi *epi, int ni, const gplum_spj_mono *spj, int nj,
                       gplum_force *force, float eps2);
/* FP64 twins: accumulate into acc[3*ni], phi[ni]; neighbour info into force (may be NULL) */
void orc_epep_f64(const gplum_epi *epi, int ni, const gplum_epj *epj, int nj,
                  double eps2, double *acc, double *phi, gplum_force *force);
void orc_epsp_quad_f64(const gplum_epi *epi, int ni, const gplum_spj_quad *spj, int nj,
                       double eps2, double *acc, double *phi);
void orc_epsp_mono_f64(const gplum_epi *epi, int ni, const gplum_spj_mono *spj, int nj,
                       double eps2, double *acc, double *phi);

/* ------------------------------------ FP64 correction + neighbour decision */
/* One (i, j) candidate of correctForceBetween2Particles[Initial]
 * (src/gravity_soft.h:75-153 / 155-242).  Returns 1 if j becomes a neighbour. */
typedef struct orc_corr_acc {
    double phi, phi_d, acc[3], acc_d[3], jerk[3], acc0;
    int32_t number;
    int64_t id_cluster;
} orc_corr_acc;
int orc_correct_pair(const gplum_params *p, const gplum_particle *pi,
                     const gplum_particle *pj, int initial, orc_corr_acc *a);

/* Neighbour graph in CSR form (directed: row i lists the j that i accepted). */
typedef struct orc_nlist {
    int32_t n;
    int32_t *offset;   /* n+1 */
    int32_t *index;    /* offset[n] entries, sorted ascending within a row */
} orc_nlist;
void orc_nlist_free(orc_nlist *nl);

/* correctForceLong / correctForceLongInitial (src/gravity_soft.h:245-372 / 374-528)
 * over the whole system.  The FP32 neighbour info (pp[i].neighbor from the soft
 * kernel) selects candidates exactly as the reference does: none if number==0,
 * id_min/id_max if number<=2, else a search within 1.1*max(r_search_i, r_search_j)
 * (the FDPS LONG_SYMMETRY neighbour query, src/particle.h:119-125).  Candidate rows
 * are processed in ascending j index.  Fills nl (may be NULL).  The Initial variant
 * also computes Sun gravity, phi_d/acc_d/jerk_d and calcDeltatInitial. */
void orc_correct_force_long(const gplum_params *p, gplum_particle *pp, int n,
                            int initial, orc_nlist *nl,
                            int32_t *n_ngb_tot, int32_t *n_with_ngb);

/* ------------------------------------------------------- cluster labelling */
/* createNeighborCluster (src/neighbor.h:799-839): id_cluster := min id over the
 * connected component of the neighbour graph. */
void orc_neighbor_cluster(gplum_particle *pp, int n, const orc_nlist *nl);
/* makeList (src/hard.h:321-438): particles with neighbours sorted by
 * (id_cluster, index); returns number of clusters, fills cl_offset[nc+1] and
 * cl_index[n_in_clusters] (caller-sized n+1 / n). */
int orc_make_cluster_list(const gplum_particle *pp, int n, int32_t *cl_offset,
                          int32_t *cl_index);

/* ----------------------------------------------------------- hard integration */
/* Sun gravity, src/gravity_hard.h:5-39. */
void orc_star_gravity(const gplum_params *p, gplum_particle *pi);
/* FPGrav::calcDeltatInitial, src/particle.h:886-914. */
void orc_calc_dt_initial(const gplum_params *p, gplum_particle *pi);
/* timeIntegrate_multi (src/hermite.h:224-479) on one cluster of n particles,
 * integrating local time 0 -> dt_tree exactly as HardSystem::timeIntegrate does
 * (FPHard(FPGrav) sets time_c=time, time=0; resetTime afterwards,
 * src/particle.h:1096-1116, 1177-1180; src/hard.h:750, 771).
 * Collisions (perfect merging) when p->collision.  Returns number of collisions;
 * *edisp receives the energy dissipated (src/collisionA.h:358-669). */
int orc_hermite_cluster(const gplum_params *p, gplum_particle *pp, int n,
                        double *edisp, int64_t *n_block_steps);
/* The same with the collision log's records (20 doubles per merger, time order; see
 * collision_record in orc_hard.c); *nrec = mergers, at most cap recorded. */
int orc_hermite_cluster_log(const gplum_params *p, gplum_particle *pp, int n, double *edisp,
                            int64_t *n_block_steps, double *rec, int cap, int *nrec);
/* Test helper: Sun + hard acc/jerk and calcDeltatInitial for one cluster. */
void orc_cluster_prepare(const gplum_params *p, gplum_particle *pp, int n);
/* timeIntegrateKepler_isolated (src/hermite.h:787-815), kepler.h:3-93. */
void orc_kepler_isolated(const gplum_params *p, gplum_particle *pi,
                         double time_start, double time_end);
/* The same with solveKeplerEq's int-abs outcome (orc_solve_kepler_int_abs). */
void orc_kepler_isolated_int_abs(const gplum_params *p, gplum_particle *pi,
                                 double time_start, double time_end);
/* timeIntegrate_isolated (src/hermite.h:734-785): Sun-only Hermite, used when
 * e >= 0.8 or eps2_sun != 0 (src/hard.h:803-811). */
void orc_hermite_isolated(const gplum_params *p, gplum_particle *pi);
/* The dispatch between the two (hard.h:803-811), n_cluster = 1. */
void orc_isolated(const gplum_params *p, gplum_particle *pi, double time_start, double time_end);
/* Kepler-equation pieces, src/kepler.h:3-93 (exposed for tests). */
double orc_cbrt_exact_ops(double x);
double orc_exp_exact_ops(double x);
double orc_solve_kepler(double l, double ecc);
/* solveKeplerEq with int abs(int) bound at kepler.h:23 (Newton only if |residual| >= 1). */
double orc_solve_kepler_int_abs(double l, double ecc);
void orc_posvel_to_elements(const double pos[3], const double vel[3], double mu,
                            double *ax, double *ecc, double *n, double *u,
                            double P[3], double Q[3]);
void orc_elements_to_posvel(double pos[3], double vel[3], double mu, double ax,
                            double ecc, double n, double u, const double P[3],
                            const double Q[3]);
double orc_eccentricity(const gplum_params *p, const gplum_particle *pi, double *ax);

/* ------------------------------------------------------ radii and energy */
/* setCutoffRadii -> calcRandomVel -> setROutRSearch (src/particle.h:1528-1543,
 * 1319-1457, 713-733). */
void orc_set_cutoff_radii(const gplum_params *p, gplum_particle *pp, int n);
/* Energy::calcEnergy (src/energy.h:29-63): out = {etot, ekin, ephi_sun, ephi, ephi_d}. */
void orc_energy(const gplum_particle *pp, int n, double out[5]);

/* ------------------------------------------------------------- octree + walk */
typedef struct orc_tree {
    int32_t n;                 /* particles                                        */
    int32_t n_node;
    double  origin[3], size;   /* root cube                                        */
    uint64_t *key;             /* sorted Morton keys, n                            */
    int32_t *perm;             /* sorted position -> input index, n                */
    int32_t *node_start, *node_count, *node_child, *node_nchild, *node_level;
    double  *node_mass, *node_com, *node_quad, *node_bmin, *node_bmax, *node_rs;
    double  *node_center;      /* geometric centre (3), half width = size/2^(level+1) */
    int32_t n_group;
    int32_t *group_node;       /* node index of each i-group                       */
    int32_t *ep_offset, *ep_leaf;   /* EP list per group = list of leaf nodes      */
    int32_t *sp_offset, *sp_node;   /* SP list per group = list of nodes           */
} orc_tree;
/* Builds the octree over pp (positions read in input order), computes moments,
 * i-groups and interaction lists.  See orc_tree.c for the construction rules. */
void orc_tree_build(const gplum_params *p, const gplum_particle *pp, int n, orc_tree *t);
void orc_tree_free(orc_tree *t);

/* Soft force over the tree lists (EP-EP + EP-SP, FP32 PIKG order), written back
 * like FDPS calcForceAllAndWriteBack + FPGrav::copyFromForce (particle.h:761-766):
 * acc, phi, neighbor (number, rank, id_max, id_min; ids are input indices).
 * use_f64: run the FP64 twins instead (acc/phi kept in double). */
int64_t orc_soft_force(const gplum_params *p, gplum_particle *pp, int n,
                       const orc_tree *t, int use_f64, int64_t *n_epsp);

/* ------------------------------------------------------------- full P3T step */
/* Main-loop state of one rank (main_p3t.cpp:241-250, read.h:111-115). */
typedef struct orc_loop {
    double r_max, r_min;        /* removeParticlesOutOfBoundary bounds (40, 0.1)           */
    int32_t reset_step;         /* periodic recompute interval (1024)                      */
    int32_t istep;              /* step counter: in before the step, incremented by it     */
    int32_t n_remove;           /* out: particles removed by the step                      */
    int32_t recomputed;         /* out: 1 if the step ended with the Initial recompute     */
    double e_now[5];            /* out: calcEnergy at the end of the step (orc_energy)     */
    gplum_particle *removed;    /* out (optional): removed particles, ascending id         */
    int32_t removed_cap;
    int32_t n_col_rec;          /* out: collision records written (cluster, then time order) */
    double *col_rec;            /* out (optional): 20 doubles per merger (orc_hermite_cluster_log) */
    int32_t col_cap;
} orc_loop;
/* Defaults of read.h:111-115 and istep = round(time_sys / dt_tree). */
void orc_loop_default(const gplum_params *p, double time_sys, orc_loop *lp);

/* Initialisation of main_p3t.cpp:289-364: radii, two rounds of tree force +
 * correctForceLongInitial; nl receives the neighbour graph. */
void orc_p3t_init(const gplum_params *p, gplum_particle *pp, int n, orc_nlist *nl,
                  double time_sys);
/* One soft step of main_p3t.cpp:423-699: kick, hard part, soft force, kick, energy,
 * MergeParticle, removeParticlesOutOfBoundary, and the recompute (setCutoffRadii + tree
 * force + correctForceLongInitial + energy) after a collision, a removal or every
 * reset_step-th step.  lp may be NULL (defaults).  Returns the number of collisions; n
 * shrinks when merged twins or out-of-bounds particles leave. */
int orc_p3t_step(const gplum_params *p, gplum_particle *pp, int *n, orc_nlist *nl,
                 double *edisp, double time_sys, orc_loop *lp);

/* ------------------------------------------------------ binary restart file */
/* makeSnapTmp / readParticleBinary (func.h:68-80; energy.h:119-138; particle.h:860-876):
 * FileHeader then one FPGrav per particle, each an fwrite of the reference's object,
 * laid out by the compiler from declarations that follow the reference's classes. */
int orc_fpgrav_size(void);
int orc_file_header_size(void);
int orc_restart_write(const char *path, int32_t id_next, double time, const double e_init[7],
                      const double e_now[7], const gplum_particle *pp, int n);
int orc_restart_read(const char *path, int32_t *n_body, int32_t *id_next, double *time, double e_init[7],
                     double e_now[7], gplum_particle *pp, int cap);

#ifdef __cplusplus
}
#endif
#endif
