/*
 * orc_hard.c -- oracle: Sun gravity, block time steps, 4th-order Hermite on a
 * cluster (with perfect-merging collisions), Kepler drift and the Sun-only
 * Hermite of isolated particles.  Test infrastructure only.
 *
 * FDPS vector expressions are evaluated left to right with scalar factors grouped
 * first (e.g. `a*b*vec` is (a*b)*vec), as the F64vec operators do.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include "oracle.h"

#define SECONDORDER  5.e-1
#define THIRDORDER   1.6666666666666667e-1
#define ALPHA_POS    1.1666666666666667e-3   /* ALPHA = 7/6000 of the position term, particle.h:339 (as written) */
#define ONE_TWELFTH  8.3333333333333333e-2
#define ONE_SIXTIETH 1.6666666666666667e-2
#define HALVE_CAP    160

static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }


/* FPHard (particle.h:970-1316) = FPGrav + the fields below. */
typedef struct orc_hp {
    gplum_particle b;
    double x0[3], v0[3], a0_s[3], j0_s[3], a0_d[3], j0_d[3], xp[3], vp[3];
    double a2_s[3], a3_s[3], a2_d[3], a3_d[3];
    double time_c;
    int32_t twin_of;   /* merge_list: index of the target this dead twin follows, or -1 */
} orc_hp;

/* ------------------------------------------------------------ Sun gravity */

/* calcStarGravity, gravity_hard.h:5-39 (phi_s too). */
void orc_star_gravity(const gplum_params *p, gplum_particle *pi)
{
    double dr[3] = {-pi->pos[0], -pi->pos[1], -pi->pos[2]};
    double dv[3] = {-pi->vel[0], -pi->vel[1], -pi->vel[2]};
    double r2inv = 1. / (dot3(dr, dr) + p->eps2_sun);
    double rinv = sqrt(r2inv);
    double r3inv = rinv * r2inv;
    double mj_rij3 = p->m_sun * r3inv;
    double alpha = dot3(dr, dv) * r2inv;
    pi->phi_s = -p->m_sun * rinv;
    for (int k = 0; k < 3; k++) {
        pi->acc_s[k] = mj_rij3 * dr[k];
        pi->jerk_s[k] = mj_rij3 * (dv[k] - 3. * alpha * dr[k]);
    }
}

/* calcStarGravity_p / _c, gravity_hard.h:41-111: same at (xp,vp) or (pos,vel), no phi. */
static void star_acc_jerk(const gplum_params *p, const double *x, const double *v,
                          double *acc_s, double *jerk_s)
{
    double dr[3] = {-x[0], -x[1], -x[2]};
    double dv[3] = {-v[0], -v[1], -v[2]};
    double r2inv = 1. / (dot3(dr, dr) + p->eps2_sun);
    double rinv = sqrt(r2inv);
    double r3inv = rinv * r2inv;
    double mj_rij3 = p->m_sun * r3inv;
    double alpha = dot3(dr, dv) * r2inv;
    for (int k = 0; k < 3; k++) {
        acc_s[k] = mj_rij3 * dr[k];
        jerk_s[k] = mj_rij3 * (dv[k] - 3. * alpha * dr[k]);
    }
}

/* ------------------------------------------------------------- time steps */

/* calcDt2nd / calcDt4th, particle.h:345-383. */
static double calc_dt2(double eta, double alpha2, double acc0, const double *acc, const double *jerk)
{
    double Acc2 = dot3(acc, acc) + alpha2 * acc0 * acc0;
    double Jerk2 = dot3(jerk, jerk);
    return (Jerk2 > 0.) ? eta * sqrt(Acc2 / Jerk2) : DBL_MAX;
}
static double calc_dt4(double eta, double alpha2, double acc0, const double *acc,
                       const double *jerk, const double *snap, const double *crac)
{
    double Acc = sqrt(dot3(acc, acc) + alpha2 * acc0 * acc0);
    double Jerk2 = dot3(jerk, jerk);
    double Jerk = sqrt(Jerk2);
    double Snap2 = dot3(snap, snap);
    double Snap = sqrt(Snap2);
    double Crac = sqrt(dot3(crac, crac));
    return (Jerk > 0.) ? eta * sqrt((Acc * Snap + Jerk2) / (Jerk * Crac + Snap2)) : DBL_MAX;
}

/* FPGrav::calcDeltatInitial, particle.h:886-914. */
void orc_calc_dt_initial(const gplum_params *p, gplum_particle *pi)
{
    double dt_next = 0.5 * p->dt_tree;
    double dt_1 = smin(calc_dt2(p->eta_0, p->alpha2, pi->acc0, pi->acc_d, pi->jerk_d),
                       calc_dt2(p->eta_sun0, p->alpha2, 0., pi->acc_s, pi->jerk_s));
    /* loops capped at HALVE_CAP iterations like the HIP code (only non-power-of-two
     * or NaN times ever reach the cap; the result is then dt_min) */
    double rem = fmod(pi->time, dt_next);
    for (int it = 0; rem != 0.0 && it < HALVE_CAP; it++) { dt_next *= 0.5; rem = fmod(pi->time, dt_next); }
    if (pi->dt > 0.)
        for (int it = 0; 2. * pi->dt < dt_next && it < HALVE_CAP; it++) dt_next *= 0.5;
    for (int it = 0; dt_1 < dt_next && it < HALVE_CAP; it++) dt_next *= 0.5;
    if (dt_next < 2. * p->dt_min || rem != 0.0) dt_next = p->dt_min;
    pi->dt = dt_next;
}

/* FPHard::calcDeltat, particle.h:1276-1314. */
static void calc_dt(const gplum_params *p, orc_hp *h)
{
    gplum_particle *b = &h->b;
    double dt_next = smin(2. * b->dt, 0.5 * p->dt_tree);
    double dt_1 = smin(calc_dt4(p->eta, p->alpha2, b->acc0, b->acc_d, b->jerk_d, h->a2_d, h->a3_d),
                       calc_dt4(p->eta_sun, p->alpha2, 0., b->acc_s, b->jerk_s, h->a2_s, h->a3_s));
    double rem = fmod(b->time, dt_next);
    for (int it = 0; rem != 0.0 && it < HALVE_CAP; it++) { dt_next *= 0.5; rem = fmod(b->time, dt_next); }
    if (rem != 0.0) { b->dt = p->dt_min; return; }
    for (int it = 0; dt_1 < dt_next && 0.5 * b->dt < dt_next && it < HALVE_CAP; it++) dt_next *= 0.5;
    if (dt_next < 2. * p->dt_min) dt_next = p->dt_min;
    b->dt = dt_next;
}

/* --------------------------------------------------------- predict / correct */

/* FPHard::predict, particle.h:1183-1207. */
static void predict(orc_hp *h, double Dt)
{
    gplum_particle *b = &h->b;
    for (int k = 0; k < 3; k++) {
        h->x0[k] = b->pos[k];
        h->v0[k] = b->vel[k];
        h->a0_s[k] = b->acc_s[k];
        h->j0_s[k] = b->jerk_s[k];
        h->j0_d[k] = b->jerk_d[k];
        h->a0_d[k] = b->acc_d[k];
        double acc_sd = b->acc_s[k] + b->acc_d[k];
        double jerk_sd = b->jerk_s[k] + b->jerk_d[k];
        h->xp[k] = b->pos[k] + Dt * (b->vel[k] + Dt * (SECONDORDER * acc_sd + THIRDORDER * Dt * jerk_sd));
        h->vp[k] = b->vel[k] + Dt * (acc_sd + Dt * SECONDORDER * jerk_sd);
    }
}

/* FPHard::correct, particle.h:1208-1274 (position term 3*ALPHA with ALPHA = 7/6000, as written). */
static void correct(orc_hp *h, double Dt)
{
    gplum_particle *b = &h->b;
    const double Dt2 = Dt * Dt;
    const double Dt_inv = 1. / Dt;
    const double Dt_inv2 = Dt_inv * Dt_inv;
    const double Dt_inv3 = Dt_inv2 * Dt_inv;
    for (int k = 0; k < 3; k++) {
        double Am_s = b->acc_s[k] - h->a0_s[k];
        double J0_s = Dt * h->j0_s[k];
        double J1_s = Dt * b->jerk_s[k];
        double A2_s = 3. * Am_s - (J1_s + 2. * J0_s);
        double A3_s = -2. * Am_s + (J1_s + J0_s);
        h->a2_s[k] = 2. * Dt_inv2 * (A2_s + 3. * A3_s);
        h->a3_s[k] = 6. * Dt_inv3 * A3_s;
        double Am_d = b->acc_d[k] - h->a0_d[k];
        double J0_d = h->j0_d[k] * Dt;
        double J1_d = b->jerk_d[k] * Dt;
        double A2_d = 3. * Am_d - (J1_d + 2. * J0_d);
        double A3_d = -2. * Am_d + (J1_d + J0_d);
        h->a2_d[k] = 2. * (A2_d + 3. * A3_d) * Dt_inv2;
        h->a3_d[k] = 6. * A3_d * Dt_inv3;
        double A2_sd = A2_s + A2_d;
        double A3_sd = A3_s + A3_d;
        b->pos[k] = h->xp[k] + ONE_SIXTIETH * Dt2 * (5. * A2_sd + 3. * ALPHA_POS * A3_sd);
        b->vel[k] = h->vp[k] + ONE_TWELFTH * Dt * (4. * A2_sd + 3. * A3_sd);
    }
}

/* ------------------------------------------------------------ pair forces */

/* Hard pair force with jerk, gravity_hard.h:219-255 (xi/vi vs xj/vj). */
static void hard_pair(const gplum_params *p, const double *xi, const double *vi, double rinv_i,
                      const double *xj, const double *vj, double massj, double rinv_j,
                      double *acc, double *jerk, double *phi)
{
    double dr[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};
    double dr2 = dot3(dr, dr);
    dr2 += p->eps2;
    double dv[3] = {vj[0] - vi[0], vj[1] - vi[1], vj[2] - vi[2]};
    double rij = sqrt(dr2);
    double rinv = 1. / rij;
    double r2inv = rinv * rinv;
    double r3inv = r2inv * rinv;
    double r_out_inv = smin(rinv_i, rinv_j);
    double mj_rij3 = massj * r3inv;
    double alpha = dot3(dr, dv) * r2inv;
    double _K = 1. - orc_cutoff_K(p, rij, r_out_inv);
    double dKdt = orc_cutoff_dKdt(p, rij, r_out_inv, alpha);
    double alpha_c = alpha * _K;
    if (phi) {
        double _W = 1. - orc_cutoff_W(p, rij, r_out_inv);
        *phi -= massj * rinv * _W;
    }
    for (int k = 0; k < 3; k++) {
        acc[k] += mj_rij3 * _K * dr[k];
        jerk[k] += mj_rij3 * (_K * dv[k] - (3. * alpha_c + dKdt) * dr[k]);
    }
}

/* calcGravity (gravity_hard.h:203-256): Sun + all other ids, with phi_d. */
static void calc_gravity_full(const gplum_params *p, orc_hp *pp, int n, int i)
{
    gplum_particle *bi = &pp[i].b;
    orc_star_gravity(p, bi);
    bi->phi_d = 0.;
    double acc[3] = {0., 0., 0.}, jerk[3] = {0., 0., 0.}, phi = 0.;
    for (int j = 0; j < n; j++) {
        if (bi->id == pp[j].b.id) continue;
        hard_pair(p, bi->pos, bi->vel, bi->r_out_inv, pp[j].b.pos, pp[j].b.vel, pp[j].b.mass,
                  pp[j].b.r_out_inv, acc, jerk, &phi);
    }
    bi->phi_d = phi;
    for (int k = 0; k < 3; k++) { bi->acc_d[k] = acc[k]; bi->jerk_d[k] = jerk[k]; }
}

/* calcGravity_p (gravity_hard.h:258-311): predicted i vs predicted j.
 * calcGravity_c (:313-367): corrected i vs predicted j. */
/* Clusters above HARD_SEQ_MAX particles sum the hard force in 64 partial sums
 * (partial l over j = l, l+64, ... ascending) added by a butterfly (partial l plus
 * partial l^d for d = 32, 16, ..., 1): the order in which the kernels evaluate one
 * particle's force with all lanes of a wave.  Smaller clusters sum sequentially in j,
 * as calcGravity_p/_c do. */
enum { HARD_SEQ_MAX = 64, HARD_LANES = 64 };

static void calc_gravity_pc(const gplum_params *p, orc_hp *pp, int n, int i, int corrected)
{
    orc_hp *h = &pp[i];
    const double *xi = corrected ? h->b.pos : h->xp;
    const double *vi = corrected ? h->b.vel : h->vp;
    star_acc_jerk(p, xi, vi, h->b.acc_s, h->b.jerk_s);
    double acc[3] = {0., 0., 0.}, jerk[3] = {0., 0., 0.};
    if (n <= HARD_SEQ_MAX) {
        for (int j = 0; j < n; j++) {
            if (h->b.id == pp[j].b.id) continue;
            hard_pair(p, xi, vi, h->b.r_out_inv, pp[j].xp, pp[j].vp, pp[j].b.mass,
                      pp[j].b.r_out_inv, acc, jerk, NULL);
        }
    } else {
        double pa[HARD_LANES][3], pj[HARD_LANES][3];
        for (int l = 0; l < HARD_LANES; l++) {
            for (int k = 0; k < 3; k++) pa[l][k] = pj[l][k] = 0.;
            for (int j = l; j < n; j += HARD_LANES) {
                if (h->b.id == pp[j].b.id) continue;
                hard_pair(p, xi, vi, h->b.r_out_inv, pp[j].xp, pp[j].vp, pp[j].b.mass,
                          pp[j].b.r_out_inv, pa[l], pj[l], NULL);
            }
        }
        for (int d = HARD_LANES / 2; d > 0; d >>= 1) {
            double qa[HARD_LANES][3], qj[HARD_LANES][3];
            for (int l = 0; l < HARD_LANES; l++)
                for (int k = 0; k < 3; k++) {
                    qa[l][k] = pa[l][k] + pa[l ^ d][k];
                    qj[l][k] = pj[l][k] + pj[l ^ d][k];
                }
            memcpy(pa, qa, sizeof(pa));
            memcpy(pj, qj, sizeof(pj));
        }
        for (int k = 0; k < 3; k++) { acc[k] = pa[0][k]; jerk[k] = pj[0][k]; }
    }
    for (int k = 0; k < 3; k++) { h->b.acc_d[k] = acc[k]; h->b.jerk_d[k] = jerk[k]; }
}

/* mergeAccJerk (hermite.h:36-74): mass-weighted acc_d/jerk_d over target + twins. */
static void merge_acc_jerk(orc_hp *pp, int n, int t)
{
    double acc[3], jerk[3], mass = pp[t].b.mass;
    for (int k = 0; k < 3; k++) { acc[k] = pp[t].b.mass * pp[t].b.acc_d[k]; jerk[k] = pp[t].b.mass * pp[t].b.jerk_d[k]; }
    for (int j = 0; j < n; j++) {
        if (pp[j].twin_of != t) continue;
        for (int k = 0; k < 3; k++) { acc[k] += pp[j].b.mass * pp[j].b.acc_d[k]; jerk[k] += pp[j].b.mass * pp[j].b.jerk_d[k]; }
        mass += pp[j].b.mass;
    }
    double mass_inv = 1. / mass;
    for (int k = 0; k < 3; k++) { acc[k] *= mass_inv; jerk[k] *= mass_inv; }
    for (int k = 0; k < 3; k++) { pp[t].b.acc_d[k] = acc[k]; pp[t].b.jerk_d[k] = jerk[k]; }
    for (int j = 0; j < n; j++) {
        if (pp[j].twin_of != t) continue;
        for (int k = 0; k < 3; k++) { pp[j].b.acc_d[k] = acc[k]; pp[j].b.jerk_d[k] = jerk[k]; }
    }
}
static void merge_acc_jerk_all(orc_hp *pp, int n)
{
    for (int t = 0; t < n; t++)
        if (pp[t].b.flags & GPLUM_FLAG_MERGED) merge_acc_jerk(pp, n, t);
}

static inline int is_dead(const orc_hp *h) { return (h->b.flags & GPLUM_FLAG_DEAD) != 0; }
static inline int is_merged(const orc_hp *h) { return (h->b.flags & GPLUM_FLAG_MERGED) != 0; }

/* collisionDetermination, hermite.h:98-196 (no MERGE_BINARY): global minimum of
 * |xp_i - xp_j| / (f_i R_i + f_j R_j) < 1 over ordered live pairs; lighter = impactor. */
static int collision_determination(orc_hp *pp, int n, int *imp, int *tar)
{
    double R_min = 1.;
    int flag = 0;
    for (int i = 0; i < n; i++) {
        if (is_dead(&pp[i])) continue;
        for (int j = 0; j < n; j++) {
            if (i == j || is_dead(&pp[j])) continue;
            double dr[3] = {pp[i].xp[0] - pp[j].xp[0], pp[i].xp[1] - pp[j].xp[1], pp[i].xp[2] - pp[j].xp[2]};
            double r1 = sqrt(dot3(dr, dr));
            double r2 = pp[i].b.f * pp[i].b.r_planet + pp[j].b.f * pp[j].b.r_planet;
            double R = r1 / r2;
            if (R < R_min) {
                R_min = R;
                if (pp[i].b.mass < pp[j].b.mass) { *imp = i; *tar = j; } else { *imp = j; *tar = i; }
                flag = 1;
            }
        }
    }
    return flag;
}

static double cutoff_W2(const gplum_params *p, double dr2, double rinv1, double rinv2)
{
    return orc_cutoff_W(p, sqrt(dr2), smin(rinv1, rinv2));
}

/* Perfect merging: Collision0::inputPair / collisionOutcome / setParticle /
 * calcEnergyDissipation (collisionA.h:91-103, 155-356, 358-669) with n_frag = 0,
 * then the merge_list update of hermite.h:413-423.  Returns edisp. */
/* Collision0::write2File's record of a merger (collisionA.h:122-148; inputPair and
 * setParticle, :155-356, with n_frag = 0), taken before perfect_merge changes the pair:
 * r[0] getTime() = time_c + time of the impactor (particle.h:1181), r[1], r[2] the ids,
 * r[3], r[4] the masses with their twins, r[5..16] pos_imp, pos_tar, vel_imp, vel_tar,
 * r[17], r[18] |pos_g_new - pos_g| / |pos_g| and the same for the velocity, where
 * pos_g_new = masspos / mass_tot with masspos = (m_imp - 0) pos_new + m_tar pos_new and
 * mass_tot = m_imp - 0 + m_tar (setParticle's sums from zero). */
enum { ORC_COLREC = 20 };
static void collision_record(const orc_hp *pp, int n, int ic, int jc, double *r)
{
    double mass_imp = pp[ic].b.mass, mass_tar = pp[jc].b.mass;
    if (is_merged(&pp[ic])) for (int a = 0; a < n; a++) if (pp[a].twin_of == ic) mass_imp += pp[a].b.mass;
    if (is_merged(&pp[jc])) for (int a = 0; a < n; a++) if (pp[a].twin_of == jc) mass_tar += pp[a].b.mass;
    double pg[3], vg[3], dp[3], dv[3];
    for (int k = 0; k < 3; k++) {
        pg[k] = (mass_imp * pp[ic].b.pos[k] + mass_tar * pp[jc].b.pos[k]) / (mass_imp + mass_tar);
        vg[k] = (mass_imp * pp[ic].b.vel[k] + mass_tar * pp[jc].b.vel[k]) / (mass_imp + mass_tar);
    }
    const double mt = mass_imp - 0. + mass_tar;
    for (int k = 0; k < 3; k++) {
        dp[k] = ((mass_imp - 0.) * pg[k] + mass_tar * pg[k]) / mt - pg[k];
        dv[k] = ((mass_imp - 0.) * vg[k] + mass_tar * vg[k]) / mt - vg[k];
    }
    r[0] = pp[ic].time_c + pp[ic].b.time;
    r[1] = (double)pp[ic].b.id;
    r[2] = (double)pp[jc].b.id;
    r[3] = mass_imp;
    r[4] = mass_tar;
    for (int k = 0; k < 3; k++) {
        r[5 + k] = pp[ic].b.pos[k]; r[8 + k] = pp[jc].b.pos[k];
        r[11 + k] = pp[ic].b.vel[k]; r[14 + k] = pp[jc].b.vel[k];
    }
    r[17] = sqrt(dot3(dp, dp) / dot3(pg, pg));
    r[18] = sqrt(dot3(dv, dv) / dot3(vg, vg));
    r[19] = 0.;
}

static double perfect_merge(const gplum_params *p, orc_hp *pp, int n, int ic, int jc)
{
    /* inputPair */
    double pos_imp[3], pos_tar[3], vel_imp[3], vel_tar[3];
    for (int k = 0; k < 3; k++) {
        pos_imp[k] = pp[ic].b.pos[k]; pos_tar[k] = pp[jc].b.pos[k];
        vel_imp[k] = pp[ic].b.vel[k]; vel_tar[k] = pp[jc].b.vel[k];
    }
    double mass_imp = pp[ic].b.mass, mass_tar = pp[jc].b.mass;
    if (is_merged(&pp[ic])) for (int a = 0; a < n; a++) if (pp[a].twin_of == ic) mass_imp += pp[a].b.mass;
    if (is_merged(&pp[jc])) for (int a = 0; a < n; a++) if (pp[a].twin_of == jc) mass_tar += pp[a].b.mass;
    double pos_g[3], vel_g[3];
    (void)pos_g;
    for (int k = 0; k < 3; k++) {
        pos_g[k] = (mass_imp * pos_imp[k] + mass_tar * pos_tar[k]) / (mass_imp + mass_tar);
        vel_g[k] = (mass_imp * vel_imp[k] + mass_tar * vel_tar[k]) / (mass_imp + mass_tar);
    }
    /* collisionOutcome */
    double pos_new[3], vel_new[3];
    for (int k = 0; k < 3; k++) {
        pos_new[k] = (mass_imp * pos_imp[k] + mass_tar * pos_tar[k]) / (mass_imp + mass_tar);
        vel_new[k] = (mass_imp * vel_imp[k] + mass_tar * vel_tar[k]) / (mass_imp + mass_tar);
    }
    /* setParticle (flag_merge, mass_frag = 0) */
    const double mass_rem = mass_imp + mass_tar;
    pp[ic].b.r_planet = 0.;
    pp[jc].b.r_planet *= orc_cbrt_exact_ops(mass_rem / mass_tar);
    const int64_t tar_id = pp[jc].b.id;
    pp[ic].b.id = tar_id;
    pp[ic].b.flags |= GPLUM_FLAG_DEAD;
    for (int a = 0; a < n; a++) if (pp[a].twin_of == ic) { pp[a].b.id = tar_id; pp[a].b.flags |= GPLUM_FLAG_DEAD; }
    pp[jc].b.flags |= GPLUM_FLAG_MERGED;
    for (int k = 0; k < 3; k++) { pp[ic].b.pos[k] = pos_new[k]; pp[ic].b.vel[k] = vel_new[k]; }
    for (int a = 0; a < n; a++) if (pp[a].twin_of == ic) for (int k = 0; k < 3; k++) { pp[a].b.pos[k] = pos_new[k]; pp[a].b.vel[k] = vel_new[k]; }
    for (int k = 0; k < 3; k++) { pp[jc].b.pos[k] = pos_new[k]; pp[jc].b.vel[k] = vel_new[k]; }
    for (int a = 0; a < n; a++) if (pp[a].twin_of == jc) for (int k = 0; k < 3; k++) { pp[a].b.pos[k] = pos_new[k]; pp[a].b.vel[k] = vel_new[k]; }

    /* calcEnergyDissipation */
    const double eps2 = p->eps2;
    const double mass_i = pp[ic].b.mass;
    /* `s * v * v` with FDPS vectors is (s*v).v */
    double e_kin = 0.;
    double vr[3], s;
    for (int k = 0; k < 3; k++) vr[k] = vel_imp[k] - vel_tar[k];
    s = 0.5 * mass_imp * mass_tar / (mass_imp + mass_tar);
    e_kin -= (s * vr[0]) * vr[0] + (s * vr[1]) * vr[1] + (s * vr[2]) * vr[2];
    for (int k = 0; k < 3; k++) vr[k] = vel_new[k] - vel_g[k];
    s = 0.5 * (mass_imp - 0.);
    e_kin += (s * vr[0]) * vr[0] + (s * vr[1]) * vr[1] + (s * vr[2]) * vr[2];
    for (int k = 0; k < 3; k++) vr[k] = vel_new[k] - vel_g[k];
    s = 0.5 * mass_tar;
    e_kin += (s * vr[0]) * vr[0] + (s * vr[1]) * vr[1] + (s * vr[2]) * vr[2];
    double e_int = 0.;
    double dr[3] = {pos_imp[0] - pos_tar[0], pos_imp[1] - pos_tar[1], pos_imp[2] - pos_tar[2]};
    double dr2 = dot3(dr, dr) + eps2;
    double rinv = sqrt(1. / dr2);
    double dphi = mass_i * rinv;
    e_int += pp[jc].b.mass * dphi;
    for (int a = 0; a < n; a++) if (pp[a].twin_of == jc) e_int += pp[a].b.mass * dphi;
    for (int i2 = 0; i2 < n; i2++) {
        if (pp[i2].twin_of != ic) continue;
        dphi = pp[i2].b.mass * rinv;
        e_int += pp[jc].b.mass * dphi;
        for (int a = 0; a < n; a++) if (pp[a].twin_of == jc) e_int += pp[a].b.mass * dphi;
    }
    (void)cutoff_W2;
    /* (id_c_frag == 0 when n_frag == 0, so the third-body loop of :562-634 is empty) */
    double e_sun = 0.;
    e_sun += mass_imp / sqrt(dot3(pos_imp, pos_imp) + eps2) + mass_tar / sqrt(dot3(pos_tar, pos_tar) + eps2);
    double d2 = dot3(pp[jc].b.pos, pp[jc].b.pos) + eps2;
    e_sun -= mass_tar * sqrt(1. / d2);
    d2 = dot3(pp[ic].b.pos, pp[ic].b.pos) + eps2;
    e_sun -= (mass_imp - 0.) * sqrt(1. / d2);
    e_sun *= p->m_sun;
    const double edisp = e_kin + e_int + e_sun;

    /* merge_list update (hermite.h:413-423) */
    pp[ic].twin_of = jc;
    if (is_merged(&pp[ic])) {
        for (int a = 0; a < n; a++) if (pp[a].twin_of == ic) pp[a].twin_of = jc;
        pp[ic].b.flags &= ~GPLUM_FLAG_MERGED;
    }
    return edisp;
}

/* timeIntegrate_multi, hermite.h:224-479, over local time [0, dt_tree]; each merger's
 * collision record (collision_record) goes to rec[0..cap) in time order (*nrec: the
 * number of mergers, rec may be NULL). */
static int hermite_cluster(const gplum_params *p, gplum_particle *part, int n, double *edisp_out,
                           int64_t *n_block_steps, double *rec, int cap, int *nrec)
{
    orc_hp *pp = (orc_hp *)calloc((size_t)n, sizeof(orc_hp));
    for (int i = 0; i < n; i++) {
        pp[i].b = part[i];
        pp[i].time_c = part[i].time;     /* FPHard(const FPGrav&) */
        pp[i].b.time = 0.;
        pp[i].twin_of = -1;
    }
    double edisp = 0.;
    int n_col = 0;
    int64_t loop = 0;
    const double time_end = p->dt_tree;
    double time = 0.;
    int *active = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));

    for (int i = 0; i < n; i++) calc_gravity_full(p, pp, n, i);
    for (int i = 0; i < n; i++) pp[i].b.flags &= ~(GPLUM_FLAG_DEAD | GPLUM_FLAG_MERGED);

    while (time < time_end) {
        /* makeActiveList */
        double t_s = 1.e10;
        int na = 0;
        for (int i = 0; i < n; i++) {
            if (is_dead(&pp[i])) continue;
            double t_i = pp[i].b.time + pp[i].b.dt;
            if (t_s > t_i) { t_s = t_i; na = 0; active[na++] = i; }
            else if (t_s == t_i) active[na++] = i;
        }
        for (int i = 0; i < n; i++) {
            for (int k = 0; k < 3; k++) { pp[i].xp[k] = 0.; pp[i].vp[k] = 0.; }
            predict(&pp[i], t_s - pp[i].b.time);
        }
        int flag_col = 0, ic = -1, jc = -1;
        if (p->collision) {
            flag_col = collision_determination(pp, n, &ic, &jc);
            if (flag_col) {
                na = 0;
                for (int i = 0; i < n; i++) {
                    pp[i].b.dt = t_s - pp[i].b.time;
                    if (is_dead(&pp[i])) continue;
                    active[na++] = i;
                }
            }
        }
        for (int a = 0; a < na; a++) {
            const int ai = active[a];
            calc_gravity_pc(p, pp, n, ai, 0);
            if (is_merged(&pp[ai])) {
                for (int j = 0; j < n; j++) if (pp[j].twin_of == ai) calc_gravity_pc(p, pp, n, j, 0);
                merge_acc_jerk(pp, n, ai);
            }
            const double t_c = t_s - pp[ai].b.time;
            for (int ite = 0; ite < 2; ite++) {
                correct(&pp[ai], t_c);
                if (is_merged(&pp[ai]))
                    for (int j = 0; j < n; j++) if (pp[j].twin_of == ai) correct(&pp[j], t_c);
                calc_gravity_pc(p, pp, n, ai, 1);
                if (is_merged(&pp[ai])) {
                    for (int j = 0; j < n; j++) if (pp[j].twin_of == ai) calc_gravity_pc(p, pp, n, j, 1);
                    merge_acc_jerk(pp, n, ai);
                }
            }
            pp[ai].b.time += t_c;
            if (is_merged(&pp[ai]))
                for (int j = 0; j < n; j++) if (pp[j].twin_of == ai) pp[j].b.time += t_c;
            if (!flag_col) {
                calc_dt(p, &pp[ai]);
                if (is_merged(&pp[ai]))
                    for (int j = 0; j < n; j++) if (pp[j].twin_of == ai) pp[j].b.dt = pp[ai].b.dt;
            }
        }
        if (flag_col) {
            if (rec && n_col < cap) collision_record(pp, n, ic, jc, rec + (size_t)n_col * ORC_COLREC);
            edisp += perfect_merge(p, pp, n, ic, jc);
            n_col++;
            for (int i = 0; i < n; i++) calc_gravity_full(p, pp, n, i);
            merge_acc_jerk_all(pp, n);
            for (int i = 0; i < n; i++) {
                if (!is_dead(&pp[i])) orc_calc_dt_initial(p, &pp[i].b);
                if (is_merged(&pp[i]))
                    for (int j = 0; j < n; j++) if (pp[j].twin_of == i) pp[j].b.dt = pp[i].b.dt;
            }
        }
        /* getSystemTime */
        double ts = 1.e10;
        for (int i = 0; i < n; i++) {
            if (is_dead(&pp[i])) continue;
            if (ts > pp[i].b.time) ts = pp[i].b.time;
        }
        time = ts;
        loop++;
    }
    for (int i = 0; i < n; i++) calc_gravity_full(p, pp, n, i);
    if (n_col > 0) merge_acc_jerk_all(pp, n);
    for (int i = 0; i < n; i++) {
        pp[i].b.n_cluster = n;
        pp[i].b.time = pp[i].time_c + pp[i].b.time;   /* resetTime */
        part[i] = pp[i].b;
    }
    free(active);
    free(pp);
    if (edisp_out) *edisp_out = edisp;
    if (n_block_steps) *n_block_steps = loop;
    if (nrec) *nrec = n_col;
    return n_col;
}

int orc_hermite_cluster(const gplum_params *p, gplum_particle *part, int n,
                        double *edisp_out, int64_t *n_block_steps)
{
    return hermite_cluster(p, part, n, edisp_out, n_block_steps, NULL, 0, NULL);
}

int orc_hermite_cluster_log(const gplum_params *p, gplum_particle *part, int n, double *edisp_out,
                            int64_t *n_block_steps, double *rec, int cap, int *nrec)
{
    return hermite_cluster(p, part, n, edisp_out, n_block_steps, rec, cap, nrec);
}

/* Test helper: state a cluster as correctForceLongInitial leaves it -- Sun and hard
 * acc/jerk (calcGravity, gravity_hard.h:203-256) then calcDeltatInitial. */
void orc_cluster_prepare(const gplum_params *p, gplum_particle *part, int n)
{
    orc_hp *pp = (orc_hp *)calloc((size_t)n, sizeof(orc_hp));
    for (int i = 0; i < n; i++) { pp[i].b = part[i]; pp[i].twin_of = -1; }
    for (int i = 0; i < n; i++) calc_gravity_full(p, pp, n, i);
    for (int i = 0; i < n; i++) {
        orc_calc_dt_initial(p, &pp[i].b);
        part[i] = pp[i].b;
    }
    free(pp);
}

/* ------------------------------------------------------------- Kepler */

/* solveKeplerEq, kepler.h:5-38 (fabs where the reference writes an unqualified abs). */
double orc_solve_kepler(double l, double ecc)
{
    double ecc2 = ecc * ecc, ecc3 = ecc2 * ecc, ecc4 = ecc2 * ecc2, ecc5 = ecc3 * ecc2, ecc6 = ecc3 * ecc3;
    double u = l
        + (ecc - ecc3 / 8. + ecc5 / 192.) * sin(l)
        + (ecc2 / 2. - ecc4 / 6. + ecc6 / 48.) * sin(2. * l)
        + (3. * ecc3 / 8. - 27. * ecc5 / 128.) * sin(3. * l)
        + (ecc4 / 3. - 4. * ecc6 / 15.) * sin(4. * l)
        + 125. * ecc5 / 384. * sin(5. * l)
        + 27. * ecc6 / 80. * sin(6. * l);
    if (fabs((u - ecc * sin(u)) - l) > 1.e-15) {
        double u0;
        int loop = 0;
        do {
            u0 = u;
            double sinu0 = sin(u0), cosu0 = cos(u0);
            u = u0 - ((u0 - ecc * sinu0 - l) / (1. - ecc * cosu0));
            loop++;
        } while (fabs(u - u0) > 1.e-15 && loop < 10);
    }
    return u;
}

/* solveKeplerEq (kepler.h:5-38) when the unqualified `abs` of kepler.h:23 binds to
 * int abs(int): the residual is truncated to an int first, so the Newton refinement
 * runs only when |u - e sin u - l| >= 1.  Which overload binds depends on the headers
 * the reference is compiled with; the kernels follow orc_solve_kepler (fabs).  This
 * variant exists so that tests can show how far the two outcomes lie apart. */
double orc_solve_kepler_int_abs(double l, double ecc)
{
    double ecc2 = ecc * ecc, ecc3 = ecc2 * ecc, ecc4 = ecc2 * ecc2, ecc5 = ecc3 * ecc2, ecc6 = ecc3 * ecc3;
    double u = l
        + (ecc - ecc3 / 8. + ecc5 / 192.) * sin(l)
        + (ecc2 / 2. - ecc4 / 6. + ecc6 / 48.) * sin(2. * l)
        + (3. * ecc3 / 8. - 27. * ecc5 / 128.) * sin(3. * l)
        + (ecc4 / 3. - 4. * ecc6 / 15.) * sin(4. * l)
        + 125. * ecc5 / 384. * sin(5. * l)
        + 27. * ecc6 / 80. * sin(6. * l);
    if (abs((int)((u - ecc * sin(u)) - l)) > 1.e-15) {
        double u0;
        int loop = 0;
        do {
            u0 = u;
            double sinu0 = sin(u0), cosu0 = cos(u0);
            u = u0 - ((u0 - ecc * sinu0 - l) / (1. - ecc * cosu0));
            loop++;
        } while (fabs(u - u0) > 1.e-15 && loop < 10);
    }
    return u;
}

/* posVel2OrbitalElement, kepler.h:39-76. */
void orc_posvel_to_elements(const double pos[3], const double vel[3], double mu,
                            double *ax, double *ecc, double *n, double *u,
                            double P[3], double Q[3])
{
    double r2 = dot3(pos, pos);
    double r = sqrt(r2);
    double rinv = 1. / r;
    double v2 = dot3(vel, vel);
    double rv = dot3(pos, vel);
    *ax = 1.0 / (2.0 * rinv - v2 / mu);
    double ecccosu = 1. - r / *ax;
    double eccsinu = rv / sqrt(mu * *ax);
    *ecc = sqrt(ecccosu * ecccosu + eccsinu * eccsinu);
    *n = sqrt(mu / (*ax * *ax * *ax));
    double cosu, sinu;
    if (*ecc != 0) {
        *u = atan2(eccsinu, ecccosu);
        cosu = ecccosu / *ecc;
        sinu = eccsinu / *ecc;
    } else {
        *u = 0.;
        cosu = 1.;
        sinu = 0.;
    }
    double aninv = sqrt(*ax / mu);
    double ecc_sq = sqrt(1. - *ecc * *ecc);
    for (int k = 0; k < 3; k++) {
        P[k] = rinv * cosu * pos[k] - aninv * sinu * vel[k];
        Q[k] = (rinv * sinu * pos[k] + aninv * (cosu - *ecc) * vel[k]) / ecc_sq;
    }
}

/* orbitalElement2PosVel, kepler.h:77-93. */
void orc_elements_to_posvel(double pos[3], double vel[3], double mu, double ax,
                            double ecc, double n, double u, const double P[3],
                            const double Q[3])
{
    (void)mu;
    double cosu = cos(u), sinu = sin(u);
    double ecc_sq = sqrt(1. - ecc * ecc);
    for (int k = 0; k < 3; k++) pos[k] = ax * ((cosu - ecc) * P[k] + ecc_sq * sinu * Q[k]);
    double rinv = sqrt(1. / dot3(pos, pos));
    for (int k = 0; k < 3; k++) vel[k] = ax * ax * n * rinv * (-sinu * P[k] + ecc_sq * cosu * Q[k]);
}

/* timeIntegrateKepler_isolated, hermite.h:787-815, with the Kepler solver given. */
static void kepler_isolated(const gplum_params *p, gplum_particle *pi, double time_start, double time_end,
                            double (*solve)(double, double))
{
    double ax, ecc, n, u, P[3], Q[3];
    orc_posvel_to_elements(pi->pos, pi->vel, p->m_sun, &ax, &ecc, &n, &u, P, Q);
    double l = u - ecc * sin(u);
    l += n * (time_end - time_start);
    u = solve(l, ecc);
    orc_elements_to_posvel(pi->pos, pi->vel, p->m_sun, ax, ecc, n, u, P, Q);
    pi->time += (time_end - time_start);
    pi->phi_d = 0.;
    for (int k = 0; k < 3; k++) { pi->acc_d[k] = 0.; pi->jerk_d[k] = 0.; }
    orc_star_gravity(p, pi);
    orc_calc_dt_initial(p, pi);
}

void orc_kepler_isolated(const gplum_params *p, gplum_particle *pi, double time_start, double time_end)
{
    kepler_isolated(p, pi, time_start, time_end, orc_solve_kepler);
}

void orc_kepler_isolated_int_abs(const gplum_params *p, gplum_particle *pi, double time_start, double time_end)
{
    kepler_isolated(p, pi, time_start, time_end, orc_solve_kepler_int_abs);
}

/* timeIntegrate_isolated, hermite.h:734-785, through FPHard (local time). */
void orc_hermite_isolated(const gplum_params *p, gplum_particle *part)
{
    orc_hp h;
    memset(&h, 0, sizeof(h));
    h.b = *part;
    h.time_c = part->time;
    h.b.time = 0.;
    orc_star_gravity(p, &h.b);
    while (h.b.time < p->dt_tree) {
        predict(&h, h.b.dt);
        for (int k = 0; k < 3; k++) { h.b.acc_d[k] = 0.; h.b.jerk_d[k] = 0.; }
        star_acc_jerk(p, h.xp, h.vp, h.b.acc_s, h.b.jerk_s);
        for (int ite = 0; ite < 2; ite++) {
            correct(&h, h.b.dt);
            star_acc_jerk(p, h.b.pos, h.b.vel, h.b.acc_s, h.b.jerk_s);
        }
        h.b.time += h.b.dt;
        calc_dt(p, &h);
    }
    h.b.phi_s = 0.;
    h.b.phi_d = 0.;
    orc_star_gravity(p, &h.b);
    h.b.time = h.time_c + h.b.time;
    *part = h.b;
}

/* The isolated-particle branch of HardSystem::timeIntegrate (hard.h:793-818): the
 * Kepler drift when e < 0.8 and the Sun is unsoftened, else the Sun-only Hermite;
 * n_cluster = 1 either way. */
void orc_isolated(const gplum_params *p, gplum_particle *pi, double time_start, double time_end)
{
    if (orc_eccentricity(p, pi, NULL) < 0.8 && p->eps2_sun == 0.)
        orc_kepler_isolated(p, pi, time_start, time_end);
    else
        orc_hermite_isolated(p, pi);
    pi->n_cluster = 1;
}
