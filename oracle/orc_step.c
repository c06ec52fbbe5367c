/*
 * orc_step.c -- oracle: initialisation and one soft step of the P3T loop
 * (main_p3t.cpp:289-364, 423-699) on one rank.  Test infrastructure only.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* tree soft force + correctForceLong[Initial] */
static void soft_and_correct(const gplum_params *p, gplum_particle *pp, int n,
                             int initial, orc_nlist *nl)
{
    orc_tree t;
    orc_tree_build(p, pp, n, &t);
    orc_soft_force(p, pp, n, &t, 0, NULL);
    orc_tree_free(&t);
    orc_nlist_free(nl);
    orc_correct_force_long(p, pp, n, initial, nl, NULL, NULL);
}

/* main_p3t.cpp:289-364 */
void orc_p3t_init(const gplum_params *p, gplum_particle *pp, int n, orc_nlist *nl, double time_sys)
{
    for (int i = 0; i < n; i++) {
        pp[i].time = time_sys;
        pp[i].neighbor.number = 0;
        pp[i].n_cluster = 0;
        pp[i].flags = 0;
        if (pp[i].r_planet <= 0.) pp[i].r_planet = pow(0.75 * pp[i].mass / (3.14159265358979323846 * p->dens), 1. / 3.);
        if (pp[i].f == 0.) pp[i].f = p->increase_factor;
    }
    orc_set_cutoff_radii(p, pp, n);
    for (int round = 0; round < 2; round++) soft_and_correct(p, pp, n, 1, nl);
}

/* MergeParticle (func.h:160-207): fold dead twins into their target, remove them. */
static int merge_particles(gplum_particle *pp, int n, double *edisp)
{
    char *rm = (char *)calloc((size_t)n + 1, 1);
    double ed = 0.;
    for (int i = 0; i < n; i++) {
        if (!((pp[i].flags & GPLUM_FLAG_MERGED) && !(pp[i].flags & GPLUM_FLAG_DEAD))) continue;
        for (int j = 0; j < n; j++) {
            if (pp[j].id == pp[i].id && i != j) {
                double mi = pp[i].mass, mj = pp[j].mass;
                double vr[3] = {pp[j].vel[0] - pp[i].vel[0], pp[j].vel[1] - pp[i].vel[1], pp[j].vel[2] - pp[i].vel[2]};
                pp[i].mass += mj;
                for (int k = 0; k < 3; k++) pp[i].vel[k] = (mi * pp[i].vel[k] + mj * pp[j].vel[k]) / (mi + mj);
                pp[i].phi = (mi * pp[i].phi + mj * pp[j].phi) / (mi + mj);
                pp[i].phi_d = (mi * pp[i].phi_d + mj * pp[j].phi_d) / (mi + mj);
                double s = 0.5 * mi * mj / (mi + mj);
                ed -= (s * vr[0]) * vr[0] + (s * vr[1]) * vr[1] + (s * vr[2]) * vr[2];
                rm[j] = 1;
            }
        }
        pp[i].flags &= ~GPLUM_FLAG_MERGED;
    }
    int m = 0;
    for (int i = 0; i < n; i++) if (!rm[i]) pp[m++] = pp[i];
    free(rm);
    *edisp += ed;
    return m;
}

static int cmp_removed(const void *a, const void *b)
{
    const gplum_particle *x = *(const gplum_particle *const *)a, *y = *(const gplum_particle *const *)b;
    return (x->id < y->id) ? -1 : (x->id > y->id);
}

/* removeParticlesOutOfBoundary (func.h:209-380) on one rank: particles with
 * |x|^2 > r_max^2 or |x|^2 < r_min^2 leave the system.  The energy they carry is booked
 * as dissipated (func.h:305-333): per removed particle i, edisp -= 1/2 m v.v, m phi_s,
 * m phi_d, m phi, and for every earlier removed j of another id edisp += -m_i m_j /
 * sqrt(|x_i - x_j|^2 + eps2).  The removed particles are taken in ascending id (the
 * reference's order is that of its OpenMP push_back into remove_list, one thread:
 * array order).  They are copied to removed[0..cap) in that order for the remove log
 * (func.h:347-352).  Returns the number removed; the survivors keep their order. */
static int remove_out_of_boundary(const gplum_params *p, gplum_particle *pp, int *n_io, double r_max, double r_min,
                                  double *edisp, gplum_particle *removed, int cap)
{
    const int n = *n_io;
    const double rmax2 = r_max * r_max, rmin2 = r_min * r_min;
    int k = 0;
    for (int i = 0; i < n; i++) {
        const double *x = pp[i].pos;
        const double pos2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
        if (pos2 > rmax2 || pos2 < rmin2) k++;
    }
    if (k == 0) return 0;
    const gplum_particle **rl = (const gplum_particle **)malloc(sizeof(*rl) * (size_t)k);
    gplum_particle *copy = (gplum_particle *)malloc(sizeof(gplum_particle) * (size_t)k);
    int a = 0, m = 0;
    for (int i = 0; i < n; i++) {
        const double *x = pp[i].pos;
        const double pos2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
        if (pos2 > rmax2 || pos2 < rmin2) copy[a++] = pp[i];
    }
    for (a = 0; a < k; a++) rl[a] = &copy[a];
    qsort(rl, (size_t)k, sizeof(*rl), cmp_removed);
    double loc = 0.;
    for (a = 0; a < k; a++) {
        const gplum_particle *q = rl[a];
        const double mi = q->mass, s = 0.5 * mi;
        loc -= (s * q->vel[0]) * q->vel[0] + (s * q->vel[1]) * q->vel[1] + (s * q->vel[2]) * q->vel[2];
        loc -= mi * q->phi_s;
        loc -= mi * q->phi_d;
        loc -= mi * q->phi;
        for (int b = 0; b < a; b++) {
            const gplum_particle *r = rl[b];
            if (q->id == r->id) continue;
            const double dr[3] = {q->pos[0] - r->pos[0], q->pos[1] - r->pos[1], q->pos[2] - r->pos[2]};
            const double rinv = 1. / sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2] + p->eps2);
            loc += -mi * r->mass * rinv;
        }
        if (removed && a < cap) removed[a] = *q;
    }
    for (int i = 0; i < n; i++) {
        const double *x = pp[i].pos;
        const double pos2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
        if (!(pos2 > rmax2 || pos2 < rmin2)) pp[m++] = pp[i];
    }
    *edisp += loc;
    *n_io = m;
    free(rl);
    free(copy);
    return k;
}

void orc_loop_default(const gplum_params *p, double time_sys, orc_loop *lp)
{
    memset(lp, 0, sizeof(*lp));
    lp->r_max = 40.;
    lp->r_min = 0.1;
    lp->reset_step = 1024;
    lp->istep = (int32_t)llround(time_sys / p->dt_tree);
}

/* One soft step, main_p3t.cpp:423-699. */
int orc_p3t_step(const gplum_params *p, gplum_particle *pp, int *n_io, orc_nlist *nl,
                 double *edisp, double time_sys, orc_loop *lp)
{
    orc_loop dflt;
    if (!lp) { orc_loop_default(p, time_sys, &dflt); lp = &dflt; }
    int n = *n_io;
    const double hdt = 0.5 * p->dt_tree;
    /* 1st kick (gravity_soft.h:5-12, particle.h:878-884) */
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) pp[i].vel[k] += hdt * pp[i].acc[k];
    /* clusters */
    orc_neighbor_cluster(pp, n, nl);
    int32_t *cl_off = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 2));
    int32_t *cl_idx = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int nc = orc_make_cluster_list(pp, n, cl_off, cl_idx);
    int n_col = 0;
    double ed = 0.;
    /* collision records per cluster, joined in cluster order afterwards */
    double **crec = lp->col_rec ? (double **)calloc((size_t)nc + 1, sizeof(double *)) : NULL;
    int *ncrec = lp->col_rec ? (int *)calloc((size_t)nc + 1, sizeof(int)) : NULL;
#pragma omp parallel for schedule(dynamic, 2) reduction(+ : n_col, ed)
    for (int c = 0; c < nc; c++) {
        int m = cl_off[c + 1] - cl_off[c];
        gplum_particle *buf = (gplum_particle *)malloc(sizeof(gplum_particle) * m);
        for (int a = 0; a < m; a++) buf[a] = pp[cl_idx[cl_off[c] + a]];
        double e = 0.;
        if (crec) {
            crec[c] = (double *)malloc(sizeof(double) * 20 * (size_t)m);
            n_col += orc_hermite_cluster_log(p, buf, m, &e, NULL, crec[c], m, &ncrec[c]);
        } else {
            n_col += orc_hermite_cluster(p, buf, m, &e, NULL);
        }
        ed += e;
        for (int a = 0; a < m; a++) pp[cl_idx[cl_off[c] + a]] = buf[a];
        free(buf);
    }
    lp->n_col_rec = 0;
    if (crec) {
        for (int c = 0; c < nc; c++) {
            for (int a = 0; a < ncrec[c] && lp->n_col_rec < lp->col_cap; a++)
                memcpy(lp->col_rec + 20 * (size_t)lp->n_col_rec++, crec[c] + 20 * (size_t)a, sizeof(double) * 20);
            free(crec[c]);
        }
        free(crec);
        free(ncrec);
    }
    free(cl_off);
    free(cl_idx);
    /* isolated particles (hard.h:793-817) */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        if (pp[i].neighbor.number) continue;
        orc_isolated(p, &pp[i], time_sys, time_sys + p->dt_tree);
    }
    *edisp += ed;
    /* soft force + correction */
    soft_and_correct(p, pp, n, 0, nl);
    /* 2nd kick */
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) pp[i].vel[k] += hdt * pp[i].acc[k];
    double e[5];
    orc_energy(pp, n, e);
    if (n_col) n = merge_particles(pp, n, edisp);
    const int n_remove = remove_out_of_boundary(p, pp, &n, lp->r_max, lp->r_min, edisp, lp->removed,
                                                lp->removed_cap);
    const int reset = lp->reset_step > 0 && lp->istep % lp->reset_step == lp->reset_step - 1;
    if (n_col || n_remove || reset) {
        orc_set_cutoff_radii(p, pp, n);
        soft_and_correct(p, pp, n, 1, nl);
        orc_energy(pp, n, e);
    }
    memcpy(lp->e_now, e, sizeof(e));
    lp->n_remove = n_remove;
    lp->recomputed = (n_col || n_remove || reset) ? 1 : 0;
    lp->istep++;
    *n_io = n;
    return n_col;
}
