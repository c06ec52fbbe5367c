/*
 * orc_io.c -- oracle: the binary restart file snap_tmp.dat (makeSnapTmp, func.h:68-80)
 * written and read as the reference does: FileHeader::writeBinary (energy.h:119-126)
 * then FPGrav::writeBinary per particle (particle.h:868-876), both an fwrite of the
 * object.  The object layouts are declared here the way the reference declares its
 * classes (EPIGrav <- EPJGrav <- FPGrav, particle.h:93-170, 404-495, default build of
 * Makefile:29-36), nested base-first, and the C compiler lays them out; the library's
 * flat record (include/gplum_system.h) is checked against these bytes.  Test
 * infrastructure only.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

typedef struct { double x, y, z; } f64vec;                 /* PS::F64vec */
typedef struct {                                             /* EPIGrav, USE_INDIVIDUAL_CUTOFF */
    int32_t id_local, myrank;
    f64vec pos;
    double r_out, r_search;
} epi_grav;
typedef struct {                                             /* EPJGrav */
    epi_grav base;
    int64_t id;
    double mass;
    f64vec vel, acc_d;
} epj_grav;
typedef struct {                                             /* NeighborInfo, particle.h:16-68 */
    int32_t number, rank, id_max, id_min;
} neighbor_info;
typedef struct {                                             /* FPGrav */
    epj_grav base;
    f64vec acc, acc_s, jerk_d, jerk_s, acc_gd;
    double phi, phi_d, phi_s;
    double v_disp;                                           /* USE_INDIVIDUAL_CUTOFF, !CONSTANT_RANDOM_VELOCITY */
    double r_out_inv;                                        /* USE_INDIVIDUAL_CUTOFF */
    double time, dt, acc0;
    double r_planet, f;
    neighbor_info neighbor;
    int64_t id_cluster;
    int32_t n_cluster;
    _Bool inDomain, isSent, isDead, isMerged;
} fp_grav;
typedef struct {                                             /* Energy, energy.h:3-11 */
    double etot, ekin, ephi_sun, ephi_planet, ephi, ephi_d, edisp;
} energy_rec;
typedef struct {                                             /* FileHeader, energy.h:70-79 */
    int32_t n_body, id_next;
    double time;
    energy_rec e_init, e_now;
} file_header;

int orc_fpgrav_size(void) { return (int)sizeof(fp_grav); }
int orc_file_header_size(void) { return (int)sizeof(file_header); }

static void vset(f64vec *v, const double a[3]) { v->x = a[0]; v->y = a[1]; v->z = a[2]; }
static void vget(const f64vec *v, double a[3]) { a[0] = v->x; a[1] = v->y; a[2] = v->z; }

int orc_restart_write(const char *path, int32_t id_next, double time, const double e_init[7],
                      const double e_now[7], const gplum_particle *pp, int n)
{
    FILE *fp = fopen(path, "wb");
    if (!fp) return -1;
    file_header h;
    memset(&h, 0, sizeof(h));
    h.n_body = n;
    h.id_next = id_next;
    h.time = time;
    memcpy(&h.e_init, e_init, sizeof(h.e_init));
    memcpy(&h.e_now, e_now, sizeof(h.e_now));
    int ok = fwrite(&h, sizeof(h), 1, fp) == 1;
    for (int i = 0; i < n && ok; i++) {
        const gplum_particle *q = &pp[i];
        fp_grav r;
        memset(&r, 0, sizeof(r));
        r.base.base.id_local = i;
        vset(&r.base.base.pos, q->pos);
        r.base.base.r_out = q->r_out;
        r.base.base.r_search = q->r_search;
        r.base.id = q->id;
        r.base.mass = q->mass;
        vset(&r.base.vel, q->vel);
        vset(&r.base.acc_d, q->acc_d);
        vset(&r.acc, q->acc);
        vset(&r.acc_s, q->acc_s);
        vset(&r.jerk_d, q->jerk_d);
        vset(&r.jerk_s, q->jerk_s);
        r.phi = q->phi; r.phi_d = q->phi_d; r.phi_s = q->phi_s;
        r.v_disp = q->v_disp; r.r_out_inv = q->r_out_inv;
        r.time = q->time; r.dt = q->dt; r.acc0 = q->acc0;
        r.r_planet = q->r_planet; r.f = q->f;
        memcpy(&r.neighbor, &q->neighbor, sizeof(r.neighbor));
        r.id_cluster = q->id_cluster;
        r.n_cluster = q->n_cluster;
        r.inDomain = 1;
        r.isDead = (q->flags & GPLUM_FLAG_DEAD) != 0;
        r.isMerged = (q->flags & GPLUM_FLAG_MERGED) != 0;
        ok = fwrite(&r, sizeof(r), 1, fp) == 1;
    }
    if (fclose(fp) != 0) ok = 0;
    return ok ? 0 : -1;
}

int orc_restart_read(const char *path, int32_t *n_body, int32_t *id_next, double *time, double e_init[7],
                     double e_now[7], gplum_particle *pp, int cap)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return -1;
    file_header h;
    int ok = fread(&h, sizeof(h), 1, fp) == 1;
    if (ok) {
        *n_body = h.n_body;
        *id_next = h.id_next;
        *time = h.time;
        memcpy(e_init, &h.e_init, sizeof(h.e_init));
        memcpy(e_now, &h.e_now, sizeof(h.e_now));
    }
    for (int i = 0; ok && pp && i < h.n_body && i < cap; i++) {
        fp_grav r;
        ok = fread(&r, sizeof(r), 1, fp) == 1;
        if (!ok) break;
        gplum_particle *q = &pp[i];
        memset(q, 0, sizeof(*q));
        q->id = r.base.id;
        q->id_cluster = r.id_cluster;
        q->mass = r.base.mass;
        q->r_planet = r.r_planet;
        q->f = r.f;
        vget(&r.base.base.pos, q->pos);
        vget(&r.base.vel, q->vel);
        vget(&r.acc, q->acc);
        vget(&r.base.acc_d, q->acc_d);
        vget(&r.jerk_d, q->jerk_d);
        vget(&r.acc_s, q->acc_s);
        vget(&r.jerk_s, q->jerk_s);
        q->phi = r.phi; q->phi_d = r.phi_d; q->phi_s = r.phi_s;
        q->r_out = r.base.base.r_out; q->r_out_inv = r.r_out_inv; q->r_search = r.base.base.r_search;
        q->v_disp = r.v_disp;
        q->time = r.time; q->dt = r.dt; q->acc0 = r.acc0;
        memcpy(&q->neighbor, &r.neighbor, sizeof(r.neighbor));
        q->n_cluster = r.n_cluster;
        q->flags = (r.isDead ? GPLUM_FLAG_DEAD : 0) | (r.isMerged ? GPLUM_FLAG_MERGED : 0);
    }
    fclose(fp);
    return ok ? 0 : -1;
}
