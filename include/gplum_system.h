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

#ifdef __cplusplus
}
#endif
#endif
