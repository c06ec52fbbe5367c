/*
 * orc_tree.c -- oracle: octree, moments, i-groups and interaction lists, then the
 * soft force over those lists.  Test infrastructure only.
 *
 * FDPS (the reference's tree library, FDPS-7.0 per the GPLUM user guide) is not
 * vendored, so this file states the tree rules the build uses; the HIP tree
 * (gplum_amd/csrc/tree.hip) implements the same rules and is checked against it.
 *
 *  keys   : root cube = particle bbox min corner, side = 1.0000001 * max extent;
 *           21-bit integer coordinates floor((x-origin)*2^21/side) interleaved as a
 *           63-bit Morton key (x bit highest within each triple).
 *  order  : particles sorted by (key, input index).
 *  nodes  : breadth-first; a node splits into its non-empty octants (ascending
 *           octant digit) when count > n_leaf_limit and level < 21.
 *  moments: FDPS MomentQuadrupole semantics (particle.h:1649-1715): mass, centre
 *           of mass, quad = sum m (x-com)(x-com)^T; internal nodes combine children
 *           with the parallel-axis term (accumulate2).  Also the particle bbox and
 *           the max r_search of the node.
 *  groups : maximal nodes with count <= n_group_limit (chunked if a max-depth leaf
 *           is larger), ordered by first particle.
 *  walk   : breadth-first from the root.  A node is opened when
 *             (a) width^2 > theta^2 * d^2, d = distance from the group bbox to the
 *                 node's centre of mass (Barnes-Hut opening), or
 *             (b) the gap between the group bbox and the node's particle bbox is
 *                 < 1.1*max(rs_group, rs_node) (SEARCH_MODE_LONG_SYMMETRY with
 *                 getRSearch() = 1.1 r_search, particle.h:119-125),
 *           opened leaves go to the EP list, unopened nodes to the SP list.
 *           (b) guarantees every pair closer than 1.1 max(r_search) - hence every
 *           changeover pair and every neighbour candidate - is an EP-EP pair.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define MAXLEVEL 21

static inline uint64_t spread3(uint64_t v)
{
    v &= 0x1fffffULL;
    v = (v | (v << 32)) & 0x1f00000000ffffULL;
    v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}

static inline int64_t coord21(double x, double o, double scale)
{
    double t = (x - o) * scale;
    int64_t c = (int64_t)floor(t);
    if (c < 0) c = 0;
    if (c > 2097151) c = 2097151;
    return c;
}

typedef struct { uint64_t key; int32_t idx; } keyidx;

static int cmp_keyidx(const void *a, const void *b)
{
    const keyidx *x = (const keyidx *)a, *y = (const keyidx *)b;
    if (x->key < y->key) return -1;
    if (x->key > y->key) return 1;
    return (x->idx < y->idx) ? -1 : (x->idx > y->idx);
}

typedef struct { int32_t *a; int64_t n, cap; } ivec;
static void ivec_push(ivec *v, int32_t x)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 1024;
        v->a = (int32_t *)realloc(v->a, sizeof(int32_t) * v->cap);
    }
    v->a[v->n++] = x;
}

void orc_tree_free(orc_tree *t)
{
    free(t->key); free(t->perm);
    free(t->node_start); free(t->node_count); free(t->node_child); free(t->node_nchild);
    free(t->node_level); free(t->node_mass); free(t->node_com); free(t->node_quad);
    free(t->node_bmin); free(t->node_bmax); free(t->node_rs); free(t->node_center);
    free(t->group_node); free(t->ep_offset); free(t->ep_leaf); free(t->sp_offset);
    free(t->sp_node);
    memset(t, 0, sizeof(*t));
}

static inline int digit_at(uint64_t key, int level)
{
    /* octant digit of the child of a level-`level` node */
    return (int)((key >> (3 * (MAXLEVEL - 1 - level))) & 7ULL);
}

void orc_tree_build(const gplum_params *p, const gplum_particle *pp, int n, orc_tree *t)
{
    memset(t, 0, sizeof(*t));
    t->n = n;
    if (n <= 0) return;
    /* bounding box */
    double mn[3], mx[3];
    for (int k = 0; k < 3; k++) { mn[k] = pp[0].pos[k]; mx[k] = pp[0].pos[k]; }
    for (int i = 1; i < n; i++)
        for (int k = 0; k < 3; k++) {
            if (pp[i].pos[k] < mn[k]) mn[k] = pp[i].pos[k];
            if (pp[i].pos[k] > mx[k]) mx[k] = pp[i].pos[k];
        }
    double ext = 0.;
    for (int k = 0; k < 3; k++) { double e = mx[k] - mn[k]; if (e > ext) ext = e; }
    if (ext == 0.) ext = 1.;
    const double side = ext * 1.0000001;
    const double scale = 2097152.0 / side;
    for (int k = 0; k < 3; k++) t->origin[k] = mn[k];
    t->size = side;

    keyidx *ki = (keyidx *)malloc(sizeof(keyidx) * n);
    for (int i = 0; i < n; i++) {
        uint64_t cx = (uint64_t)coord21(pp[i].pos[0], mn[0], scale);
        uint64_t cy = (uint64_t)coord21(pp[i].pos[1], mn[1], scale);
        uint64_t cz = (uint64_t)coord21(pp[i].pos[2], mn[2], scale);
        ki[i].key = (spread3(cx) << 2) | (spread3(cy) << 1) | spread3(cz);
        ki[i].idx = i;
    }
    qsort(ki, n, sizeof(keyidx), cmp_keyidx);
    t->key = (uint64_t *)malloc(sizeof(uint64_t) * n);
    t->perm = (int32_t *)malloc(sizeof(int32_t) * n);
    for (int i = 0; i < n; i++) { t->key[i] = ki[i].key; t->perm[i] = ki[i].idx; }
    free(ki);

    /* breadth-first node construction */
    int64_t cap = 2 * (int64_t)n + 16;
    t->node_start = (int32_t *)malloc(sizeof(int32_t) * cap);
    t->node_count = (int32_t *)malloc(sizeof(int32_t) * cap);
    t->node_child = (int32_t *)malloc(sizeof(int32_t) * cap);
    t->node_nchild = (int32_t *)malloc(sizeof(int32_t) * cap);
    t->node_level = (int32_t *)malloc(sizeof(int32_t) * cap);
    int32_t nn = 1;
    t->node_start[0] = 0; t->node_count[0] = n; t->node_level[0] = 0;
    for (int32_t c = 0; c < nn; c++) {
        const int32_t s = t->node_start[c], cnt = t->node_count[c], lv = t->node_level[c];
        t->node_child[c] = -1;
        t->node_nchild[c] = 0;
        if (cnt <= p->n_leaf_limit || lv >= MAXLEVEL) continue;
        int32_t b = s;
        const int32_t e = s + cnt;
        while (b < e) {
            int d = digit_at(t->key[b], lv);
            int32_t q = b;
            while (q < e && digit_at(t->key[q], lv) == d) q++;
            if (nn >= cap) {
                cap *= 2;
                t->node_start = (int32_t *)realloc(t->node_start, sizeof(int32_t) * cap);
                t->node_count = (int32_t *)realloc(t->node_count, sizeof(int32_t) * cap);
                t->node_child = (int32_t *)realloc(t->node_child, sizeof(int32_t) * cap);
                t->node_nchild = (int32_t *)realloc(t->node_nchild, sizeof(int32_t) * cap);
                t->node_level = (int32_t *)realloc(t->node_level, sizeof(int32_t) * cap);
            }
            if (t->node_nchild[c] == 0) t->node_child[c] = nn;
            t->node_nchild[c]++;
            t->node_start[nn] = b;
            t->node_count[nn] = q - b;
            t->node_level[nn] = lv + 1;
            nn++;
            b = q;
        }
    }
    t->n_node = nn;

    /* moments, deepest nodes first (breadth-first order reversed) */
    t->node_mass = (double *)malloc(sizeof(double) * nn);
    t->node_com = (double *)malloc(sizeof(double) * nn * 3);
    t->node_quad = (double *)malloc(sizeof(double) * nn * 6);
    t->node_bmin = (double *)malloc(sizeof(double) * nn * 3);
    t->node_bmax = (double *)malloc(sizeof(double) * nn * 3);
    t->node_rs = (double *)malloc(sizeof(double) * nn);
    t->node_center = (double *)malloc(sizeof(double) * nn * 3);
    for (int32_t c = nn - 1; c >= 0; c--) {
        double m = 0., mp[3] = {0., 0., 0.}, q[6] = {0., 0., 0., 0., 0., 0.};
        double bmn[3], bmx[3], rs = 0.;
        const int32_t s = t->node_start[c], cnt = t->node_count[c];
        if (t->node_nchild[c] == 0) {
            const gplum_particle *p0 = &pp[t->perm[s]];
            for (int k = 0; k < 3; k++) { bmn[k] = p0->pos[k]; bmx[k] = p0->pos[k]; }
            for (int32_t a = s; a < s + cnt; a++) {
                const gplum_particle *pa = &pp[t->perm[a]];
                m += pa->mass;
                for (int k = 0; k < 3; k++) {
                    mp[k] += pa->mass * pa->pos[k];
                    if (pa->pos[k] < bmn[k]) bmn[k] = pa->pos[k];
                    if (pa->pos[k] > bmx[k]) bmx[k] = pa->pos[k];
                }
                if (pa->r_search > rs) rs = pa->r_search;
            }
            double com[3];
            for (int k = 0; k < 3; k++) com[k] = (m != 0.) ? mp[k] / m : 0.;
            for (int32_t a = s; a < s + cnt; a++) {
                const gplum_particle *pa = &pp[t->perm[a]];
                double dx = pa->pos[0] - com[0], dy = pa->pos[1] - com[1], dz = pa->pos[2] - com[2];
                double cx = pa->mass * dx, cy = pa->mass * dy, cz = pa->mass * dz;
                q[0] += cx * dx; q[1] += cy * dy; q[2] += cz * dz;
                q[3] += cx * dy; q[4] += cx * dz; q[5] += cy * dz;
            }
            for (int k = 0; k < 3; k++) t->node_com[3 * c + k] = com[k];
        } else {
            const int32_t c0 = t->node_child[c], nc = t->node_nchild[c];
            for (int k = 0; k < 3; k++) { bmn[k] = t->node_bmin[3 * c0 + k]; bmx[k] = t->node_bmax[3 * c0 + k]; }
            for (int32_t ch = c0; ch < c0 + nc; ch++) {
                m += t->node_mass[ch];
                for (int k = 0; k < 3; k++) {
                    mp[k] += t->node_mass[ch] * t->node_com[3 * ch + k];
                    if (t->node_bmin[3 * ch + k] < bmn[k]) bmn[k] = t->node_bmin[3 * ch + k];
                    if (t->node_bmax[3 * ch + k] > bmx[k]) bmx[k] = t->node_bmax[3 * ch + k];
                }
                if (t->node_rs[ch] > rs) rs = t->node_rs[ch];
            }
            double com[3];
            for (int k = 0; k < 3; k++) com[k] = (m != 0.) ? mp[k] / m : 0.;
            for (int32_t ch = c0; ch < c0 + nc; ch++) {
                const double mt = t->node_mass[ch];
                double dx = t->node_com[3 * ch + 0] - com[0];
                double dy = t->node_com[3 * ch + 1] - com[1];
                double dz = t->node_com[3 * ch + 2] - com[2];
                double cx = mt * dx, cy = mt * dy, cz = mt * dz;
                const double *qc = &t->node_quad[6 * ch];
                q[0] += cx * dx + qc[0]; q[1] += cy * dy + qc[1]; q[2] += cz * dz + qc[2];
                q[3] += cx * dy + qc[3]; q[4] += cx * dz + qc[4]; q[5] += cy * dz + qc[5];
            }
            for (int k = 0; k < 3; k++) t->node_com[3 * c + k] = com[k];
        }
        t->node_mass[c] = m;
        for (int k = 0; k < 6; k++) t->node_quad[6 * c + k] = q[k];
        for (int k = 0; k < 3; k++) { t->node_bmin[3 * c + k] = bmn[k]; t->node_bmax[3 * c + k] = bmx[k]; }
        t->node_rs[c] = rs;
        /* geometric centre from the key prefix */
        const int lv = t->node_level[c];
        const uint64_t key = t->key[s];
        const double w = side / (double)(1ULL << lv);
        uint64_t ix = 0, iy = 0, iz = 0;
        for (int b = 0; b < lv; b++) {
            int d = digit_at(key, b);
            ix = (ix << 1) | ((d >> 2) & 1);
            iy = (iy << 1) | ((d >> 1) & 1);
            iz = (iz << 1) | (d & 1);
        }
        t->node_center[3 * c + 0] = mn[0] + ((double)ix + 0.5) * w;
        t->node_center[3 * c + 1] = mn[1] + ((double)iy + 0.5) * w;
        t->node_center[3 * c + 2] = mn[2] + ((double)iz + 0.5) * w;
    }

    /* i-groups: maximal nodes with count <= n_group_limit (node, start, count) */
    ivec gn = {0}, gs = {0}, gc = {0};
    {
        ivec stack = {0};
        ivec_push(&stack, 0);
        while (stack.n) {
            int32_t c = stack.a[--stack.n];
            const int32_t cnt = t->node_count[c];
            if (cnt <= p->n_group_limit || t->node_nchild[c] == 0) {
                for (int32_t off = 0; off < cnt; off += p->n_group_limit) {
                    ivec_push(&gn, c);
                    ivec_push(&gs, t->node_start[c] + off);
                    int32_t r = cnt - off;
                    ivec_push(&gc, r < p->n_group_limit ? r : p->n_group_limit);
                }
            } else {
                for (int32_t ch = t->node_child[c] + t->node_nchild[c] - 1; ch >= t->node_child[c]; ch--)
                    ivec_push(&stack, ch);
            }
        }
        free(stack.a);
    }
    /* depth-first with children pushed in reverse visits groups in particle order */
    t->n_group = (int32_t)gn.n;
    t->group_node = (int32_t *)malloc(sizeof(int32_t) * (gn.n * 3 + 1));
    for (int64_t g = 0; g < gn.n; g++) {
        t->group_node[3 * g + 0] = gn.a[g];
        t->group_node[3 * g + 1] = gs.a[g];
        t->group_node[3 * g + 2] = gc.a[g];
    }
    free(gn.a); free(gs.a); free(gc.a);

    /* walk */
    const double theta2 = p->theta * p->theta;
    ivec ep = {0}, sp = {0};
    t->ep_offset = (int32_t *)malloc(sizeof(int32_t) * (t->n_group + 1));
    t->sp_offset = (int32_t *)malloc(sizeof(int32_t) * (t->n_group + 1));
    int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * nn);
    for (int32_t g = 0; g < t->n_group; g++) {
        const int32_t s = t->group_node[3 * g + 1], cnt = t->group_node[3 * g + 2];
        double gmn[3], gmx[3], grs = 0.;
        for (int k = 0; k < 3; k++) { gmn[k] = pp[t->perm[s]].pos[k]; gmx[k] = gmn[k]; }
        for (int32_t a = s; a < s + cnt; a++) {
            const gplum_particle *pa = &pp[t->perm[a]];
            for (int k = 0; k < 3; k++) {
                if (pa->pos[k] < gmn[k]) gmn[k] = pa->pos[k];
                if (pa->pos[k] > gmx[k]) gmx[k] = pa->pos[k];
            }
            if (pa->r_search > grs) grs = pa->r_search;
        }
        t->ep_offset[g] = (int32_t)ep.n;
        t->sp_offset[g] = (int32_t)sp.n;
        int32_t qh = 0, qt = 0;
        queue[qt++] = 0;
        while (qh < qt) {
            const int32_t c = queue[qh++];
            double d2 = 0., gap2 = 0.;
            for (int k = 0; k < 3; k++) {
                const double cm = t->node_com[3 * c + k];
                double dd = 0.;
                if (cm < gmn[k]) dd = gmn[k] - cm;
                else if (cm > gmx[k]) dd = cm - gmx[k];
                d2 += dd * dd;
                double gg = 0.;
                if (t->node_bmax[3 * c + k] < gmn[k]) gg = gmn[k] - t->node_bmax[3 * c + k];
                else if (t->node_bmin[3 * c + k] > gmx[k]) gg = t->node_bmin[3 * c + k] - gmx[k];
                gap2 += gg * gg;
            }
            const double w = side / (double)(1ULL << t->node_level[c]);
            const double rsm = 1.1 * (grs > t->node_rs[c] ? grs : t->node_rs[c]);
            const int open = (w * w > theta2 * d2) || (gap2 < rsm * rsm);
            if (open) {
                if (t->node_nchild[c] == 0) {
                    ivec_push(&ep, c);
                } else {
                    for (int32_t ch = t->node_child[c]; ch < t->node_child[c] + t->node_nchild[c]; ch++)
                        queue[qt++] = ch;
                }
            } else {
                ivec_push(&sp, c);
            }
        }
    }
    t->ep_offset[t->n_group] = (int32_t)ep.n;
    t->sp_offset[t->n_group] = (int32_t)sp.n;
    t->ep_leaf = ep.a;
    t->sp_node = sp.a;
    free(queue);
}

/* Soft force over the tree lists.  For each i-group the EPI array is the group's
 * particles, the EPJ array the particles of its EP leaves (list order), the SPJ
 * array its SP nodes; the FP32 kernels accumulate into a cleared Force and the
 * result is written back as FPGrav::copyFromForce does (particle.h:761-766).
 * Returns the number of EP-EP pair evaluations. */
int64_t orc_soft_force(const gplum_params *p, gplum_particle *pp, int n,
                       const orc_tree *t, int use_f64, int64_t *n_epsp)
{
    int64_t n_epep_tot = 0, n_epsp_tot = 0;
    const float eps2 = (float)p->eps2;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : n_epep_tot, n_epsp_tot)
    for (int32_t g = 0; g < t->n_group; g++) {
        const int32_t s = t->group_node[3 * g + 1], ni = t->group_node[3 * g + 2];
        int32_t nj = 0;
        for (int32_t e = t->ep_offset[g]; e < t->ep_offset[g + 1]; e++) nj += t->node_count[t->ep_leaf[e]];
        const int32_t ns = t->sp_offset[g + 1] - t->sp_offset[g];
        gplum_epi *epi = (gplum_epi *)malloc(sizeof(gplum_epi) * ni);
        gplum_epj *epj = (gplum_epj *)malloc(sizeof(gplum_epj) * (nj > 0 ? nj : 1));
        gplum_force *f = (gplum_force *)malloc(sizeof(gplum_force) * ni);
        for (int32_t a = 0; a < ni; a++) {
            const int32_t idx = t->perm[s + a];
            epi[a].id_local = idx;
            epi[a].myrank = 0;
            for (int k = 0; k < 3; k++) epi[a].pos[k] = pp[idx].pos[k];
            epi[a].r_out = pp[idx].r_out;
            epi[a].r_search = pp[idx].r_search;
        }
        int32_t jj = 0;
        for (int32_t e = t->ep_offset[g]; e < t->ep_offset[g + 1]; e++) {
            const int32_t c = t->ep_leaf[e];
            for (int32_t a = t->node_start[c]; a < t->node_start[c] + t->node_count[c]; a++) {
                const int32_t idx = t->perm[a];
                gplum_epj *q = &epj[jj++];
                q->id_local = idx; q->myrank = 0;
                for (int k = 0; k < 3; k++) {
                    q->pos[k] = pp[idx].pos[k];
                    q->vel[k] = pp[idx].vel[k];
                    q->acc_d[k] = pp[idx].acc_d[k];
                }
                q->r_out = pp[idx].r_out; q->r_search = pp[idx].r_search;
                q->id = pp[idx].id; q->mass = pp[idx].mass;
            }
        }
        orc_force_clear(f, ni);
        if (!use_f64) {
            orc_epep_f32(epi, ni, epj, nj, f, eps2);
        } else {
            double *acc = (double *)calloc(3 * (size_t)ni + ni, sizeof(double));
            double *phi = acc + 3 * ni;
            orc_epep_f64(epi, ni, epj, nj, (double)eps2, acc, phi, f);
            if (p->quadrupole) {
                gplum_spj_quad *spj = (gplum_spj_quad *)malloc(sizeof(gplum_spj_quad) * (ns > 0 ? ns : 1));
                for (int32_t a = 0; a < ns; a++) {
                    const int32_t c = t->sp_node[t->sp_offset[g] + a];
                    spj[a].mass = t->node_mass[c];
                    for (int k = 0; k < 3; k++) spj[a].pos[k] = t->node_com[3 * c + k];
                    for (int k = 0; k < 6; k++) spj[a].quad[k] = t->node_quad[6 * c + k];
                }
                orc_epsp_quad_f64(epi, ni, spj, ns, (double)eps2, acc, phi);
                free(spj);
            } else {
                gplum_spj_mono *spj = (gplum_spj_mono *)malloc(sizeof(gplum_spj_mono) * (ns > 0 ? ns : 1));
                for (int32_t a = 0; a < ns; a++) {
                    const int32_t c = t->sp_node[t->sp_offset[g] + a];
                    spj[a].mass = t->node_mass[c];
                    for (int k = 0; k < 3; k++) spj[a].pos[k] = t->node_com[3 * c + k];
                }
                orc_epsp_mono_f64(epi, ni, spj, ns, (double)eps2, acc, phi);
                free(spj);
            }
            for (int32_t a = 0; a < ni; a++) {
                const int32_t idx = t->perm[s + a];
                for (int k = 0; k < 3; k++) pp[idx].acc[k] = acc[3 * a + k];
                pp[idx].phi = phi[a];
                pp[idx].neighbor = f[a].neighbor;
            }
            free(acc);
        }
        if (!use_f64) {
            if (p->quadrupole) {
                gplum_spj_quad *spj = (gplum_spj_quad *)malloc(sizeof(gplum_spj_quad) * (ns > 0 ? ns : 1));
                for (int32_t a = 0; a < ns; a++) {
                    const int32_t c = t->sp_node[t->sp_offset[g] + a];
                    spj[a].mass = t->node_mass[c];
                    for (int k = 0; k < 3; k++) spj[a].pos[k] = t->node_com[3 * c + k];
                    for (int k = 0; k < 6; k++) spj[a].quad[k] = t->node_quad[6 * c + k];
                }
                orc_epsp_quad_f32(epi, ni, spj, ns, f, eps2);
                free(spj);
            } else {
                gplum_spj_mono *spj = (gplum_spj_mono *)malloc(sizeof(gplum_spj_mono) * (ns > 0 ? ns : 1));
                for (int32_t a = 0; a < ns; a++) {
                    const int32_t c = t->sp_node[t->sp_offset[g] + a];
                    spj[a].mass = t->node_mass[c];
                    for (int k = 0; k < 3; k++) spj[a].pos[k] = t->node_com[3 * c + k];
                }
                orc_epsp_mono_f32(epi, ni, spj, ns, f, eps2);
                free(spj);
            }
            for (int32_t a = 0; a < ni; a++) {
                const int32_t idx = t->perm[s + a];
                for (int k = 0; k < 3; k++) pp[idx].acc[k] = (double)f[a].acc[k];
                pp[idx].phi = (double)f[a].phi;
                pp[idx].neighbor = f[a].neighbor;
            }
        }
        n_epep_tot += (int64_t)ni * nj;
        n_epsp_tot += (int64_t)ni * ns;
        free(epi); free(epj); free(f);
    }
    if (n_epsp) *n_epsp = n_epsp_tot;
    (void)n;
    return n_epep_tot;
}
