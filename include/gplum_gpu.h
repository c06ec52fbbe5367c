/*
 * gplum_gpu.h -- C ABI of the MI355X (gfx950) implementation of GPLUM's P3T hot path:
 * FP32 soft force (EP-EP changeover + neighbour counting, EP-SP monopole/quadrupole),
 * FP64 neighbour decision / changeover correction / cluster labelling, and the FP64
 * block-timestep Hermite integration of the resulting hard clusters.
 *
 * Every entry point returns an int status: 0 on success, < 0 on error
 * (GPLUM_E_*); gplum_last_error() returns a message for the calling thread.
 * No entry point retains a caller pointer after it returns.
 *
 * Struct layouts are byte-compatible with the reference's `individual_cutoff=yes`
 * build (src/particle.h:16-170, FDPS SPJMonopole/SPJQuadrupole), so FDPS arrays can
 * be passed unchanged.
 */
#ifndef GPLUM_GPU_H
#define GPLUM_GPU_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */
#define GPLUM_OK            0
#define GPLUM_E_HIP        -1   /* HIP runtime error                                  */
#define GPLUM_E_ARG        -2   /* invalid argument (null pointer, bad size)          */
#define GPLUM_E_OVERFLOW   -3   /* a device buffer was too small; caller may retry    */
#define GPLUM_E_NODEVICE   -4   /* no usable gfx950 device                             */
#define GPLUM_E_STATE      -5   /* call order violated (e.g. retrieve before dispatch) */

/* ------------------------------------------------------- reference-layout PODs */

/* NeighborInfo, src/particle.h:16-68 (16 B). Clear value: {0, 0, -1, 2^31-1}. */
typedef struct gplum_neighbor_info {
    int32_t number;
    int32_t rank;
    int32_t id_max;
    int32_t id_min;
} gplum_neighbor_info;

/* ForceGrav, src/particle.h:70-86 (32 B). */
typedef struct gplum_force {
    float acc[3];
    float phi;
    gplum_neighbor_info neighbor;
} gplum_force;

/* EPIGrav with USE_INDIVIDUAL_CUTOFF, src/particle.h:93-141 (48 B). */
typedef struct gplum_epi {
    int32_t id_local;
    int32_t myrank;
    double  pos[3];
    double  r_out;
    double  r_search;
} gplum_epi;

/* EPJGrav, src/particle.h:149-170 (112 B). */
typedef struct gplum_epj {
    int32_t id_local;
    int32_t myrank;
    double  pos[3];
    double  r_out;
    double  r_search;
    int64_t id;
    double  mass;
    double  vel[3];
    double  acc_d[3];
} gplum_epj;

/* FDPS SPJMonopole (MySPJMonopole, src/particle.h:1628-1647) (32 B). */
typedef struct gplum_spj_mono {
    double mass;
    double pos[3];
} gplum_spj_mono;

/* FDPS SPJQuadrupole (MySPJQuadrupole, src/particle.h:1717-1737) (80 B).
 * quad = sum m (x-com)(x-com)^T, stored {xx, yy, zz, xy, xz, yz}. */
typedef struct gplum_spj_quad {
    double mass;
    double pos[3];
    double quad[6];
} gplum_spj_quad;

/* --------------------------------------------------------------- run parameters */

/* The static FPGrav parameters (src/particle.h:421-479, defaults :917-961) plus the
 * tree parameters of main_p3t.cpp:345.  gplum_params_default() fills the reference
 * defaults; gplum_params_set_gamma() mirrors FPGrav::setGamma (particle.h:619-633). */
typedef struct gplum_params {
    double m_sun, dens, eps2, eps2_sun;
    double R_cut0, R_cut1, R_search0, R_search1, R_search2, R_search3;
    double gamma, g_1_inv, g_1_inv7, w_y, f1;
    double r_cut_min, r_cut_max, p_cut, increase_factor;
    double dt_tree, dt_min, eta, eta_0, eta_sun, eta_sun0, alpha2;
    double theta;
    int32_t n_leaf_limit;
    int32_t n_group_limit;
    int32_t quadrupole;      /* 1: EP-SP quadrupole (USE_QUAD), 0: monopole          */
    int32_t collision;       /* 1: perfect-merging collisions (COLLISION)            */
} gplum_params;

void gplum_params_default(gplum_params *p);
void gplum_params_set_gamma(gplum_params *p, double g);

/* ------------------------------------------------------------------- utilities */
const char *gplum_last_error(void);
int gplum_device_info(char *name, int name_len, int *n_cu, double *clock_ghz);

/* ---------------------------------------------- PIKG-shaped soft-force kernels */

/* Replaces the PIKG functor CalcForceLongEPEP(eps2)(epi, ni, epj, nj, force)
 * called from calcForceEPEPWithSearch::operator() (src/gravity_kernel.hpp:12-23;
 * DSL src/gravity_kernel_epep.pikg:1-97).  Accumulates into force[0..ni) exactly as
 * the PIKG kernel does: acc += sum m r/r^3, phi -= sum m/r with r^2 clamped at
 * max(r_out_i, r_out_j)^2, and neighbour number/rank/id_max/id_min for
 * |r|^2+eps2 < 1.0201 max(r_search)^2.  Host pointers; synchronous. */
int gplum_calc_force_epep(const gplum_epi *epi, int ni,
                          const gplum_epj *epj, int nj,
                          gplum_force *force, float eps2);

/* Replaces CalcForceLongEPSP(eps2)(epi, ni, spj, nj, force) (src/gravity_kernel.hpp:
 * 128-136; DSL src/gravity_kernel_epsp.pikg:1-100), quadrupole variant (USE_QUAD). */
int gplum_calc_force_epsp_quad(const gplum_epi *epi, int ni,
                               const gplum_spj_quad *spj, int nj,
                               gplum_force *force, float eps2);

/* Monopole variant of the same functor (USE_QUAD undefined). */
int gplum_calc_force_epsp_mono(const gplum_epi *epi, int ni,
                               const gplum_spj_mono *spj, int nj,
                               gplum_force *force, float eps2);

/* FP64 twins of the three kernels above (same arithmetic in double, results
 * rounded to the FP32 Force fields only at the end; the acc/phi sums are also
 * returned unrounded in acc64[3*ni] / phi64[ni] when those are non-null).
 * They are the "vs FP64" reference for the FP32 kernels. */
int gplum_calc_force_epep_f64(const gplum_epi *epi, int ni,
                              const gplum_epj *epj, int nj,
                              gplum_force *force, float eps2,
                              double *acc64, double *phi64);
int gplum_calc_force_epsp_quad_f64(const gplum_epi *epi, int ni,
                                   const gplum_spj_quad *spj, int nj,
                                   gplum_force *force, float eps2,
                                   double *acc64, double *phi64);

/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): gplum_calc_force_epep on device-resident arrays (d_epi[ni], d_epj[nj],
 * d_force[ni], accumulated as there), launched asynchronously on `stream` (a
 * hipStream_t; NULL = the default stream).  It allocates nothing: its i-chunk table
 * lives in a fixed per-device array of 4096 entries, so ni <= 1,048,576 (larger ni
 * returns GPLUM_E_ARG).  Every call on a device shares that array: calls that overlap
 * on one device (different streams or host threads) race. */
int gplum_calc_force_epep_noalloc(const gplum_epi *d_epi, int ni,
                                  const gplum_epj *d_epj, int nj,
                                  gplum_force *d_force, float eps2, void *stream);

/* ------------------------------------------------- FDPS MultiWalk interface */

/* eps2 used by the MultiWalk kernels (the PIKG functors take it as constructor
 * argument: CalcForceLongEPEP(FP_t::eps2), src/gravity_kernel.hpp:17-22). */
int gplum_multiwalk_set_eps2(float eps2);

/* Dispatch functor of PS::TreeForForce::calcForceAllAndWriteBackMultiWalk
 * (called at the sites of calcForceAllAndWriteBack, src/main_p3t.cpp:351-356,
 * 583-588, 672-677): n_walk i-groups with their EP and SP (quadrupole) lists are
 * packed, uploaded and launched as one EP-EP + EP-SP kernel on the tag's stream;
 * returns without waiting.  Same argument meaning as the FDPS dispatch signature
 * S32 dispatch(S32 tag, S32 n_walk, const EPI**, const S32*, const EPJ**,
 * const S32*, const SPJ**, const S32*). */
int gplum_multiwalk_dispatch(int tag, int n_walk, const gplum_epi **epi, const int *n_epi,
                             const gplum_epj **epj, const int *n_epj,
                             const gplum_spj_quad **spj, const int *n_spj);

/* The same dispatch with monopole super-particles (USE_QUAD off,
 * PS::SPJMonopole: mass, pos). */
int gplum_multiwalk_dispatch_mono(int tag, int n_walk, const gplum_epi **epi, const int *n_epi,
                                  const gplum_epj **epj, const int *n_epj,
                                  const gplum_spj_mono **spj, const int *n_spj);

/* Dispatch functor of calcForceAllAndWriteBackMultiWalkIndex: the walks' EP and SP
 * lists are index lists id_epj[w][0..n_epj[w]) / id_spj[w][0..n_spj[w]) into the tree's
 * whole arrays epj[0..n_epj_tot) / spj[0..n_spj_tot), which FDPS passes with send_flag
 * set on the first dispatch of a force call (they are uploaded then and reused by the
 * later dispatches of that call, of any tag).  Same argument meaning as the FDPS
 * signature S32 dispatch(S32 tag, S32 n_walk, const EPI**, const S32*, const S32**,
 * const S32*, const S32**, const S32*, const EPJ*, const S32, const SPJ*, const S32,
 * const bool); indices outside the whole arrays are an argument error. */
int gplum_multiwalk_dispatch_index(int tag, int n_walk, const gplum_epi **epi, const int *n_epi,
                                   const int32_t **id_epj, const int *n_epj, const int32_t **id_spj,
                                   const int *n_spj, const gplum_epj *epj, int n_epj_tot,
                                   const gplum_spj_quad *spj, int n_spj_tot, int send_flag);
int gplum_multiwalk_dispatch_index_mono(int tag, int n_walk, const gplum_epi **epi, const int *n_epi,
                                        const int32_t **id_epj, const int *n_epj, const int32_t **id_spj,
                                        const int *n_spj, const gplum_epj *epj, int n_epj_tot,
                                        const gplum_spj_mono *spj, int n_spj_tot, int send_flag);

/* Retrieve functor: waits for the tag's dispatch and assigns force[w][0..ni[w])
 * (values start from ForceGrav::clear(), src/particle.h:81-85).  FDPS signature
 * S32 retrieve(S32 tag, S32 n_walk, const S32* ni, Force** force). */
int gplum_multiwalk_retrieve(int tag, int n_walk, const int *ni, gplum_force **force);

#ifdef __cplusplus
}
#endif

#endif /* GPLUM_GPU_H */
