/*
 * orc_correct.c -- oracle: correctForceLong[Initial], neighbour query, cluster
 * labelling and segmentation, cut-off radii and energy.  Test infrastructure only.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include "oracle.h"

static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }

/* --------------------------------------------------------- neighbour query */

/* Replacement of FDPS getNeighborListOneParticle (gravity_soft.h:296-297): all
 * j != i with |x_i - x_j| < 1.1 max(r_search_i, r_search_j).  Any superset of
 * the pairs closer than max(r_search) yields the same corrected result, because
 * the FP64 decision and the changeover correction (r < r_out <= r_search) are
 * applied per candidate.  Implemented with a sorted cell hash. */
typedef struct { uint64_t key; int32_t idx; } cellkey;
static int cmp_cellkey(const void *a, const void *b)
{
    const cellkey *x = (const cellkey *)a, *y = (const cellkey *)b;
    if (x->key < y->key) return -1;
    if (x->key > y->key) return 1;
    return (x->idx < y->idx) ? -1 : (x->idx > y->idx);
}
typedef struct { cellkey *ck; int n; double h, o[3]; } cellgrid;

static uint64_t cell_of(const cellgrid *g, const double *x, int d[3])
{
    int64_t c[3];
    for (int k = 0; k < 3; k++) {
        c[k] = (int64_t)floor((x[k] - g->o[k]) / g->h) + (d ? d[k] : 0);
        if (c[k] < 0) c[k] = 0;
        if (c[k] > 2097151) c[k] = 2097151;
    }
    return ((uint64_t)c[0] << 42) | ((uint64_t)c[1] << 21) | (uint64_t)c[2];
}

static void grid_build(cellgrid *g, const gplum_particle *pp, int n)
{
    double rs = 0., mn[3] = {0., 0., 0.};
    for (int i = 0; i < n; i++) {
        if (pp[i].r_search > rs) rs = pp[i].r_search;
        for (int k = 0; k < 3; k++) if (i == 0 || pp[i].pos[k] < mn[k]) mn[k] = pp[i].pos[k];
    }
    g->h = 1.1 * rs * 1.0001;
    if (g->h <= 0.) g->h = 1.;
    for (int k = 0; k < 3; k++) g->o[k] = mn[k] - 2. * g->h;
    g->n = n;
    g->ck = (cellkey *)malloc(sizeof(cellkey) * (n > 0 ? n : 1));
    for (int i = 0; i < n; i++) { g->ck[i].key = cell_of(g, pp[i].pos, NULL); g->ck[i].idx = i; }
    qsort(g->ck, n, sizeof(cellkey), cmp_cellkey);
}

static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x < y) ? -1 : (x > y);
}

/* returns number of candidates written to out (capacity cap), ascending index */
static int grid_query(const cellgrid *g, const gplum_particle *pp, int i, int32_t **out, int *cap)
{
    int m = 0;
    for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dz = -1; dz <= 1; dz++) {
                int d[3] = {dx, dy, dz};
                uint64_t key = cell_of(g, pp[i].pos, d);
                int lo = 0, hi = g->n;
                while (lo < hi) { int mid = (lo + hi) / 2; if (g->ck[mid].key < key) lo = mid + 1; else hi = mid; }
                for (int a = lo; a < g->n && g->ck[a].key == key; a++) {
                    int j = g->ck[a].idx;
                    if (j == i) continue;
                    double rx = pp[j].pos[0] - pp[i].pos[0];
                    double ry = pp[j].pos[1] - pp[i].pos[1];
                    double rz = pp[j].pos[2] - pp[i].pos[2];
                    double r2 = rx * rx + ry * ry + rz * rz;
                    double rs = 1.1 * smax(pp[i].r_search, pp[j].r_search);
                    if (r2 < rs * rs) {
                        if (m == *cap) { *cap = *cap ? 2 * *cap : 64; *out = (int32_t *)realloc(*out, sizeof(int32_t) * *cap); }
                        (*out)[m++] = j;
                    }
                }
            }
    qsort(*out, m, sizeof(int32_t), cmp_i32);
    return m;
}

/* ------------------------------------------------------- correctForceLong */

void orc_star_gravity(const gplum_params *p, gplum_particle *pi);
void orc_calc_dt_initial(const gplum_params *p, gplum_particle *pi);

void orc_correct_force_long(const gplum_params *p, gplum_particle *pp, int n,
                            int initial, orc_nlist *nl,
                            int32_t *n_ngb_tot, int32_t *n_with_ngb)
{
    cellgrid g;
    grid_build(&g, pp, n);
    /* old acc_d of every particle feeds the re-search filter (gravity_soft.h:388-392) */
    double *acc_d_new = initial ? (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1)) : NULL;
    gplum_particle *old = (gplum_particle *)malloc(sizeof(gplum_particle) * (n > 0 ? n : 1));
    memcpy(old, pp, sizeof(gplum_particle) * n);
    int32_t *cnt = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    int32_t **rows = (int32_t **)calloc((size_t)n + 1, sizeof(int32_t *));
    int64_t tot = 0, with = 0;

#pragma omp parallel for schedule(dynamic, 256) reduction(+ : tot, with)
    for (int i = 0; i < n; i++) {
        gplum_particle *pi = &pp[i];
        if (initial) orc_star_gravity(p, pi);
        orc_corr_acc a;
        memset(&a, 0, sizeof(a));
        a.id_cluster = pi->id;
        const int32_t number = old[i].neighbor.number;
        a.phi += pi->mass * pi->r_out_inv;   /* self term, gravity_soft.h:280/292 */
        int32_t *cand = NULL;
        int cap = 0, nc = 0;
        if (number > 0) {
            if (number > 2 || old[i].neighbor.rank != 0) {
                nc = grid_query(&g, old, i, &cand, &cap);
            } else {
                cap = 2;
                cand = (int32_t *)malloc(sizeof(int32_t) * 2);
                cand[nc++] = old[i].neighbor.id_min;          /* getId(0) */
                if (number == 2) cand[nc++] = old[i].neighbor.id_max; /* getId(1) */
            }
        }
        int32_t *acc_rows = NULL;
        int nacc = 0;
        if (nc > 0) acc_rows = (int32_t *)malloc(sizeof(int32_t) * nc);
        for (int c = 0; c < nc; c++) {
            const int j = cand[c];
            if (old[j].id == old[i].id) continue;
            /* pi fields used: pos/vel/acc_d (old values), r_out/r_out_inv/r_search */
            if (orc_correct_pair(p, &old[i], &old[j], initial, &a)) acc_rows[nacc++] = j;
        }
        free(cand);
        rows[i] = acc_rows;
        cnt[i] = nacc;
        pi->neighbor.number = a.number;
        pi->id_cluster = a.id_cluster;
        if (a.number) { tot += a.number; with++; }
        for (int k = 0; k < 3; k++) pi->acc[k] += a.acc[k];
        pi->phi += a.phi;
        pi->acc0 = (a.acc0 > 0.) ? (double)a.number / a.acc0 : 0.;
        if (initial) {
            for (int k = 0; k < 3; k++) { acc_d_new[3 * i + k] = a.acc_d[k]; pi->jerk_d[k] = a.jerk[k]; }
            pi->phi_d = a.phi_d;
            /* calcDeltatInitial runs before acc_d is written back
             * (gravity_soft.h:502-516): it sees the new jerk_d, acc0 and Sun terms
             * but the previous acc_d. */
            orc_calc_dt_initial(p, pi);
        }
    }
    if (initial) {
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) pp[i].acc_d[k] = acc_d_new[3 * i + k];
        free(acc_d_new);
    }
    if (nl) {
        nl->n = n;
        nl->offset = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
        nl->offset[0] = 0;
        for (int i = 0; i < n; i++) nl->offset[i + 1] = nl->offset[i] + cnt[i];
        nl->index = (int32_t *)malloc(sizeof(int32_t) * ((size_t)nl->offset[n] + 1));
        for (int i = 0; i < n; i++)
            if (cnt[i]) memcpy(&nl->index[nl->offset[i]], rows[i], sizeof(int32_t) * cnt[i]);
    }
    for (int i = 0; i < n; i++) free(rows[i]);
    free(rows);
    free(cnt);
    free(old);
    free(g.ck);
    if (n_ngb_tot) *n_ngb_tot = (int32_t)tot;
    if (n_with_ngb) *n_with_ngb = (int32_t)with;
}

/* ------------------------------------------------------------- clusters */

/* createNeighborCluster (neighbor.h:799-839): repeat "both labels := min" over
 * the local pairs (i < j) until nothing changes.  The fixed point is the minimum
 * id over each connected component; computed here by the same sweep. */
void orc_neighbor_cluster(gplum_particle *pp, int n, const orc_nlist *nl)
{
    int changed = 1;
    while (changed) {
        changed = 0;
        for (int i = 0; i < n; i++)
            for (int32_t e = nl->offset[i]; e < nl->offset[i + 1]; e++) {
                int j = nl->index[e];
                if (!(i < j)) continue;   /* pair_list holds i < j_id_local */
                if (pp[i].id_cluster != pp[j].id_cluster) {
                    int64_t m = pp[i].id_cluster < pp[j].id_cluster ? pp[i].id_cluster : pp[j].id_cluster;
                    pp[i].id_cluster = pp[j].id_cluster = m;
                    changed = 1;
                }
            }
    }
}

typedef struct { int64_t cid, id; int32_t idx; } cidkey;
static int cmp_cid(const void *a, const void *b)
{
    const cidkey *x = (const cidkey *)a, *y = (const cidkey *)b;
    if (x->cid < y->cid) return -1;
    if (x->cid > y->cid) return 1;
    if (x->id < y->id) return -1;
    if (x->id > y->id) return 1;
    return (x->idx < y->idx) ? -1 : (x->idx > y->idx);
}

/* HardSystem::makeList (hard.h:321-438) on one rank: particles with neighbours,
 * grouped by id_cluster.  The reference lists a cluster's members in particle-array
 * order, which on one rank is the order of the initial conditions (ascending id)
 * until FDPS removes merged particles; that order decides the merge target between
 * equal masses (collisionDetermination, hermite.h:116-139) and the FP64 summation
 * order inside the cluster.  Here members are listed by ascending id, so the result
 * does not depend on how particles are laid out in memory. */
int orc_make_cluster_list(const gplum_particle *pp, int n, int32_t *cl_offset, int32_t *cl_index)
{
    cidkey *k = (cidkey *)malloc(sizeof(cidkey) * (n > 0 ? n : 1));
    int nk = 0;
    for (int i = 0; i < n; i++)
        if (pp[i].neighbor.number) {
            k[nk].cid = pp[i].id_cluster;
            k[nk].id = pp[i].id;
            k[nk].idx = i;
            nk++;
        }
    qsort(k, nk, sizeof(cidkey), cmp_cid);
    int nc = 0;
    for (int a = 0; a < nk; a++) {
        if (a == 0 || k[a].cid != k[a - 1].cid) cl_offset[nc++] = a;
        cl_index[a] = k[a].idx;
    }
    cl_offset[nc] = nk;
    free(k);
    return nc;
}

/* ----------------------------------------------------------------- radii */

/* FPGrav::getEccentricity (particle.h:668-681); *ax receives the semi-major axis. */
double orc_eccentricity(const gplum_params *p, const gplum_particle *pi, double *ax)
{
    const double *x = pi->pos, *v = pi->vel;
    double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    double rv = x[0] * v[0] + x[1] * v[1] + x[2] * v[2];
    double a = 1.0 / (2.0 / r - (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / p->m_sun);
    double ecccosu = 1. - r / a;
    double eccsinu2 = rv * rv / (p->m_sun * a);
    if (ax) *ax = a;
    return sqrt(ecccosu * ecccosu + eccsinu2);
}

typedef struct { int64_t id; int32_t idx; } idkey;
static int cmp_id(const void *a, const void *b)
{
    const idkey *x = (const idkey *)a, *y = (const idkey *)b;
    return (x->id < y->id) ? -1 : (x->id > y->id);
}

/* calcRandomVel (non-ISOTROPIC branch, particle.h:1319-1457) then setROutRSearch
 * (particle.h:713-733).  The bin sums are taken in ascending particle-id order, in
 * chunks of 64 (the reference's OpenMP reduction order is unspecified; a fixed order
 * makes the radii reproducible across particle orderings and lets the kernels sum the
 * chunks in parallel with the same result).  exp and the cube root are the
 * exact-operation versions of orc_math.c, shared bit for bit with the kernels. */
void orc_set_cutoff_radii(const gplum_params *p, gplum_particle *pp, int n)
{
    double r_max = 0., r_min = -DBL_MAX;
    for (int i = 0; i < n; i++) {
        double r2 = pp[i].pos[0] * pp[i].pos[0] + pp[i].pos[1] * pp[i].pos[1];
        r_max = smax(r_max, r2);
        r_min = smax(r_min, -r2);
    }
    r_max = sqrt(r_max) * 1.01;
    r_min = sqrt(-r_min) * 0.99;
    enum { NB = 32 };
    const double dr = (r_max - r_min) / NB;
    const double drinv = 1. / dr;
    /* bin sums in a fixed blocked order: chunks of 64 consecutive values (ascending
     * id) summed left to right, then the chunk sums left to right (the kernels sum
     * the same way, in parallel over chunks) */
    enum { CHUNK = 64 };
    double v_disp[NB], chunk[NB];
    int32_t n_ptcl[NB], in_chunk[NB];
    for (int b = 0; b < NB; b++) { v_disp[b] = 0.; n_ptcl[b] = 0; chunk[b] = 0.; in_chunk[b] = 0; }
    idkey *ord = (idkey *)malloc(sizeof(idkey) * (n > 0 ? n : 1));
    for (int i = 0; i < n; i++) { ord[i].id = pp[i].id; ord[i].idx = i; }
    qsort(ord, n, sizeof(idkey), cmp_id);
    for (int a = 0; a < n; a++) {
        const gplum_particle *pi = &pp[ord[a].idx];
        double r2 = pi->pos[0] * pi->pos[0] + pi->pos[1] * pi->pos[1];
        double ri = sqrt(r2);
        int j = (int)((ri - r_min) * drinv);
        if (j == NB) j = NB - 1;
        double vk = sqrt(p->m_sun / sqrt(pi->pos[0] * pi->pos[0] + pi->pos[1] * pi->pos[1]));
        double vkx = -pi->pos[1] / ri, vky = pi->pos[0] / ri, vkz = 0.;
        vkx *= vk; vky *= vk; vkz *= vk;
        double vx = pi->vel[0] - vkx, vy = pi->vel[1] - vky, vz = pi->vel[2] - vkz;
        chunk[j] += vx * vx + vy * vy + vz * vz;
        n_ptcl[j]++;
        if (++in_chunk[j] == CHUNK) { v_disp[j] += chunk[j]; chunk[j] = 0.; in_chunk[j] = 0; }
    }
    for (int b = 0; b < NB; b++) if (in_chunk[b]) v_disp[b] += chunk[b];
    free(ord);
    for (int b = 0; b < NB; b++) v_disp[b] = (n_ptcl[b] > 0) ? sqrt(v_disp[b] / n_ptcl[b]) : 0.;
    for (int i = 0; i < n; i++) {
        gplum_particle *pi = &pp[i];
        double r2 = pi->pos[0] * pi->pos[0] + pi->pos[1] * pi->pos[1];
        double vd = 0., ni = 0.;
        for (int b = 0; b < NB; b++) {
            double ddr = r_min + (b + 0.5) * dr - sqrt(r2);
            double expr = orc_exp_exact_ops(-ddr * ddr * drinv * drinv);
            vd += v_disp[b] * expr;
            ni += (n_ptcl[b] > 0) ? expr : 0.;
        }
        pi->v_disp = vd / ni;
        /* setROutRSearch */
        double ax;
        double ecc = orc_eccentricity(p, pi, &ax);
        double ax2 = (ecc < 0.6) ? ax : sqrt(pi->pos[0] * pi->pos[0] + pi->pos[1] * pi->pos[1] + pi->pos[2] * pi->pos[2]);
        double rHill = orc_cbrt_exact_ops(pi->mass / (3. * p->m_sun)) * ax2;
        /* pow(ax2, -p_cut) is exactly 1 for the default p_cut = 0 (library pow may
         * differ by an ulp between host and device otherwise) */
        double pc = (p->p_cut == 0.) ? 1. : pow(ax2, -p->p_cut);
        double r_out_i = smax(p->R_cut0 * pc * rHill, p->R_cut1 * pi->v_disp * p->dt_tree);
        double r_out = smax(r_out_i, p->r_cut_min);
        if (p->r_cut_max > 0.) r_out = smin(p->r_cut_max, r_out);
        pi->r_out = r_out;
        pi->r_out_inv = 1. / r_out;
        pi->r_search = p->R_search0 * r_out + p->R_search1 * pi->v_disp * p->dt_tree;
    }
}

/* Energy::calcEnergy, energy.h:29-63. out = {etot, ekin, ephi_sun, ephi, ephi_d} */
void orc_energy(const gplum_particle *pp, int n, double out[5])
{
    double ekin = 0., esun = 0., ephi = 0., ephid = 0.;
    for (int i = 0; i < n; i++) {
        const gplum_particle *q = &pp[i];
        /* mass * vel * vel with FDPS vectors is (mass*vel).vel */
        ekin += (q->mass * q->vel[0]) * q->vel[0] + (q->mass * q->vel[1]) * q->vel[1]
                + (q->mass * q->vel[2]) * q->vel[2];
        esun += q->mass * q->phi_s;
        ephi += q->mass * q->phi;
        ephid += q->mass * q->phi_d;
    }
    ekin *= 0.5;
    ephi *= 0.5;
    ephid *= 0.5;
    out[1] = ekin;
    out[2] = esun;
    out[3] = ephi;
    out[4] = ephid;
    out[0] = ekin + esun + (ephi + ephid);
}
