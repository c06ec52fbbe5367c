/*
 * orc_soft.c -- oracle: changeover functions, FP32/FP64 soft kernels and the FP64
 * changeover correction / neighbour decision.  Test infrastructure only.
 */
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include <float.h>
#include "oracle.h"

/* std::min / std::max semantics (NaN handling matters, see orc_correct_pair). */
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }

/* ---------------------------------------------------------------- parameters */

/* Reference defaults: static member initialisers of src/particle.h:917-961,
 * plus tree defaults (src/read.h: theta 0.5, n_leaf_limit 8, n_group_limit 256). */
void orc_params_default(gplum_params *p)
{
    memset(p, 0, sizeof(*p));
    p->m_sun = 1.;
    p->dens = 5.049667e6;
    p->eps2 = 0.;
    p->eps2_sun = 0.;
    p->R_cut0 = 2.;
    p->R_cut1 = 8.;
    p->R_search0 = 1.;
    p->R_search1 = 4.;
    p->R_search2 = 1.;
    p->R_search3 = 4.;
    p->r_cut_min = 0.;
    p->r_cut_max = 0.;
    p->p_cut = 0.;
    p->increase_factor = 1.;
    p->dt_tree = 1.0 / 32.0;         /* pow2(-5) */
    p->dt_min = 1.0 / 8192.0;        /* pow2(-13) */
    p->eta = 0.01;
    p->eta_0 = 0.001;
    p->eta_sun = 0.01;
    p->eta_sun0 = 0.001;
    p->alpha2 = 1.;
    p->theta = 0.5;
    p->n_leaf_limit = 8;
    p->n_group_limit = 256;
    p->quadrupole = 1;
    p->collision = 1;
    orc_params_set_gamma(p, 0.5);
}

/* FPGrav::setGamma, src/particle.h:619-633. */
void orc_params_set_gamma(gplum_params *p, double g)
{
    p->gamma = g;
    p->g_1_inv = 1. / (g - 1.);
    double g2 = g * g;
    double g_1_inv3 = p->g_1_inv * p->g_1_inv * p->g_1_inv;
    p->g_1_inv7 = g_1_inv3 * g_1_inv3 * p->g_1_inv;
    p->w_y = 7. / 3. * ((((((g - 9.) * g + 45.) * g - 60. * log(g)) * g - 45.) * g + 9.) * g - 1.)
             * p->g_1_inv7;
    p->f1 = (-10. / 3. + 14. * (g + 1.) - 21. * ((g + 3.) * g + 1.)
             + 35. / 3. * (((g + 9.) * g + 9.) * g + 1.)
             - 70. * ((g + 3.) * g + 1.) * g
             + 210. * (g + 1.) * g2
             + (((g - 7.) * g + 21.) * g - 35.) * g2 * g2) * p->g_1_inv7;
}

/* ---------------------------------------------------------- src/cutfunc.h */

/* cutoff_f, src/cutfunc.h:4-13. */
double orc_cutoff_f(const gplum_params *p, double y)
{
    const double g = p->gamma;
    const double g2 = g * g;
    return (((((((-10. / 3. * y + 14. * (g + 1.)) * y - 21. * ((g + 3.) * g + 1.)) * y
                + 35. / 3. * (((g + 9.) * g + 9.) * g + 1.)) * y
               - 70. * ((g + 3.) * g + 1.) * g) * y
              + 210. * (g + 1.) * g2) * y - 140. * g2 * g * log(y)) * y
            + (((g - 7.) * g + 21.) * g - 35.) * g2 * g2) * p->g_1_inv7;
}

/* cutoff_W, src/cutfunc.h:16-28. */
double orc_cutoff_W(const gplum_params *p, double rij, double r_out_inv)
{
    const double g = p->gamma;
    const double y = rij * r_out_inv;
    if (1.0 <= y) {
        return 1.0;
    } else if (y <= g) {
        return y * p->w_y;
    } else {
        double f = orc_cutoff_f(p, y);
        return f + y * (1. - p->f1);
    }
}

/* cutoff_K(x), src/cutfunc.h:30-39. */
double orc_cutoff_Kx(double x)
{
    if (x < 0.) return 0.;
    if (x >= 1.) return 1.;
    double x2 = x * x;
    return (((-20. * x + 70.) * x - 84.) * x + 35.) * x2 * x2;
}

/* cutoff_dKdr, src/cutfunc.h:40-45. */
double orc_cutoff_dKdr(const gplum_params *p, double x, double r_out_inv)
{
    const double x_1 = x - 1.;
    return (x < 0. || x >= 1.) ? 0.
           : (140. * x * x * x * x_1 * x_1 * x_1 * r_out_inv * p->g_1_inv);
}

/* cutoff_K(rij, r_out_inv), src/cutfunc.h:59-62. */
double orc_cutoff_K(const gplum_params *p, double rij, double r_out_inv)
{
    const double x = (p->gamma - rij * r_out_inv) * p->g_1_inv;
    return orc_cutoff_Kx(x);
}

/* cutoff_dKdt, src/cutfunc.h:63-66. */
double orc_cutoff_dKdt(const gplum_params *p, double rij, double r_out_inv, double alpha)
{
    const double x = (p->gamma - rij * r_out_inv) * p->g_1_inv;
    return alpha * rij * orc_cutoff_dKdr(p, x, r_out_inv);
}

/* --------------------------------------------------------- FP32 soft kernels */

void orc_force_clear(gplum_force *f, int n)
{
    /* ForceGrav::clear / NeighborInfo::clear, src/particle.h:48-53, 81-85. */
    for (int i = 0; i < n; i++) {
        f[i].acc[0] = f[i].acc[1] = f[i].acc[2] = 0.f;
        f[i].phi = 0.f;
        f[i].neighbor.number = 0;
        f[i].neighbor.rank = 0;
        f[i].neighbor.id_max = -1;
        f[i].neighbor.id_min = 2147483647;  /* S32_MAX-1 with S32_MAX = 2^31 (main_p3t.cpp:32) */
    }
}

/*
 * EP-EP FP32 kernel with neighbour counting: gravity_kernel_epep.pikg:53-97
 * (accumulation convention of the PIKG functor: af += m r/r^3, pf -= m/r,
 * counts accumulated, id_max/id_min max/min-merged; rank += drank^2, the PIKG form).
 *
 * Fixed FP32 operation order (the HIP kernel uses the same):
 *   xiloc = (float)(xi - xi0)            (difference in double, rounded once)
 *   xjloc = (float)(xj - xi0)
 *   rij   = xjloc - xiloc
 *   r2_real = fma(rz, rz, fma(ry, ry, fma(rx, rx, eps2)))
 *   rout2    = max(rout_i^2, rout_j^2)           (rout^2 per particle, f32)
 *   rsearch2 = max(s_i, s_j), s = (rs*rs)*1.0201f (f32)
 *   r2 = max(r2_real, rout2)
 *   neighbour iff r2_real < rsearch2 && (idi != idj || ranki != rankj)
 * max(a,b)^2 == max(a^2,b^2) bit-for-bit because rounding is monotonic, so the
 * per-particle squares equal the reference's max-then-square.
 * The acc/phi sums use 1/sqrtf + one Newton step like the DSL (:332-334); their
 * last bits are not part of the parity contract (compared with a tolerance).
 */
void orc_epep_f32(const gplum_epi *epi, int ni, const gplum_epj *epj, int nj,
                  gplum_force *force, float eps2)
{
    if (ni <= 0) return;
    const double x0 = epi[0].pos[0], y0 = epi[0].pos[1], z0 = epi[0].pos[2];
    float *xj = (float *)malloc(sizeof(float) * (size_t)(nj > 0 ? nj : 1) * 6);
    float *yj = xj + nj, *zj = yj + nj, *mj = zj + nj, *ro2j = mj + nj, *rs2j = ro2j + nj;
    for (int j = 0; j < nj; j++) {
        xj[j] = (float)(epj[j].pos[0] - x0);
        yj[j] = (float)(epj[j].pos[1] - y0);
        zj[j] = (float)(epj[j].pos[2] - z0);
        mj[j] = (float)epj[j].mass;
        float ro = (float)epj[j].r_out, rs = (float)epj[j].r_search;
        ro2j[j] = ro * ro;
        rs2j[j] = (rs * rs) * 1.0201f;
    }
    for (int i = 0; i < ni; i++) {
        const float xi = (float)(epi[i].pos[0] - x0);
        const float yi = (float)(epi[i].pos[1] - y0);
        const float zi = (float)(epi[i].pos[2] - z0);
        const float roi = (float)epi[i].r_out, rsi = (float)epi[i].r_search;
        const float ro2i = roi * roi, rs2i = (rsi * rsi) * 1.0201f;
        const int idi = epi[i].id_local, ranki = epi[i].myrank;
        float ax = 0.f, ay = 0.f, az = 0.f, pot = 0.f;
        int ngb = 0, rankf = 0, idmax = -1, idmin = 2147483647;
        for (int j = 0; j < nj; j++) {
            const float rx = xj[j] - xi, ry = yj[j] - yi, rz = zj[j] - zi;
            const float r2_real = fmaf(rz, rz, fmaf(ry, ry, fmaf(rx, rx, eps2)));
            const float rout2 = ro2i > ro2j[j] ? ro2i : ro2j[j];
            const float rsearch2 = rs2i > rs2j[j] ? rs2i : rs2j[j];
            const float r2 = r2_real > rout2 ? r2_real : rout2;
            if (r2_real < rsearch2) {
                if (idi != epj[j].id_local || ranki != epj[j].myrank) {
                    int rsub = ranki - epj[j].myrank;
                    ngb += 1;
                    rankf += rsub * rsub;
                    if (epj[j].id_local > idmax) idmax = epj[j].id_local;
                    if (epj[j].id_local < idmin) idmin = epj[j].id_local;
                }
            }
            float r_inv = 1.0f / sqrtf(r2);
            const float tmp = 3.0f - r2 * (r_inv * r_inv);
            r_inv *= (tmp * 0.5f);
            const float r2_inv = r_inv * r_inv;
            const float mr_inv = mj[j] * r_inv;
            const float mr3_inv = r2_inv * mr_inv;
            ax += mr3_inv * rx;
            ay += mr3_inv * ry;
            az += mr3_inv * rz;
            pot -= mr_inv;
        }
        force[i].acc[0] += ax;
        force[i].acc[1] += ay;
        force[i].acc[2] += az;
        force[i].phi += pot;
        force[i].neighbor.number += ngb;
        force[i].neighbor.rank += rankf;
        if (idmax > force[i].neighbor.id_max) force[i].neighbor.id_max = idmax;
        if (idmin < force[i].neighbor.id_min) force[i].neighbor.id_min = idmin;
    }
    free(xj);
}

/* EP-SP quadrupole FP32, gravity_kernel_epsp.pikg:47-100 with the correct trace
 * tr = xx+yy+zz (:75; the non-PIKG functor's xx+yy+xx trace, gravity_kernel.hpp:177, is NOT
 * reproduced). */
void orc_epsp_quad_f32(const gplum_epi *epi, int ni, const gplum_spj_quad *spj, int nj,
                       gplum_force *force, float eps2)
{
    if (ni <= 0) return;
    const double x0 = epi[0].pos[0], y0 = epi[0].pos[1], z0 = epi[0].pos[2];
    for (int i = 0; i < ni; i++) {
        const float xi = (float)(epi[i].pos[0] - x0);
        const float yi = (float)(epi[i].pos[1] - y0);
        const float zi = (float)(epi[i].pos[2] - z0);
        float ax = 0.f, ay = 0.f, az = 0.f, pot = 0.f;
        for (int j = 0; j < nj; j++) {
            const float xj = (float)(spj[j].pos[0] - x0);
            const float yj = (float)(spj[j].pos[1] - y0);
            const float zj = (float)(spj[j].pos[2] - z0);
            const float m = (float)spj[j].mass;
            const float qxx0 = (float)spj[j].quad[0], qyy0 = (float)spj[j].quad[1];
            const float qzz0 = (float)spj[j].quad[2], qxy0 = (float)spj[j].quad[3];
            const float qxz0 = (float)spj[j].quad[4], qyz0 = (float)spj[j].quad[5];
            const float rx = xj - xi, ry = yj - yi, rz = zj - zi;
            const float r2 = fmaf(rz, rz, fmaf(ry, ry, fmaf(rx, rx, eps2)));
            float r_inv = 1.0f / sqrtf(r2);
            const float tmp = 3.0f - r2 * (r_inv * r_inv);
            r_inv *= (tmp * 0.5f);
            const float r2_inv = r_inv * r_inv;
            const float r3_inv = r2_inv * r_inv;
            const float r4_inv = r2_inv * r2_inv;
            const float r5_inv = r2_inv * r3_inv;
            const float tr = qxx0 + qyy0 + qzz0;
            const float qxx = 3.0f * qxx0 - tr;
            const float qyy = 3.0f * qyy0 - tr;
            const float qzz = 3.0f * qzz0 - tr;
            const float qxy = 3.0f * qxy0;
            const float qyz = 3.0f * qyz0;
            const float qzx = 3.0f * qxz0;
            const float mtr = -(eps2 * tr);
            const float qrx = qxx * rx + qxy * ry + qzx * rz;
            const float qry = qyy * ry + qyz * rz + qxy * rx;
            const float qrz = qzz * rz + qzx * rx + qyz * ry;
            const float rqr = mtr + qrx * rx + qry * ry + qrz * rz;
            const float rqr_r4_inv = rqr * r4_inv;
            const float meff = m + 0.5f * rqr_r4_inv;
            const float meff3 = (m + 2.5f * rqr_r4_inv) * r3_inv;
            pot -= meff * r_inv;
            ax = ax - r5_inv * qrx + meff3 * rx;
            ay = ay - r5_inv * qry + meff3 * ry;
            az = az - r5_inv * qrz + meff3 * rz;
        }
        force[i].acc[0] += ax;
        force[i].acc[1] += ay;
        force[i].acc[2] += az;
        force[i].phi += pot;
    }
}

/* EP-SP monopole FP32, gravity_kernel_epsp.pikg:103-108 (#else branch). */
void orc_epsp_mono_f32(const gplum_epi *epi, int ni, const gplum_spj_mono *spj, int nj,
                       gplum_force *force, float eps2)
{
    if (ni <= 0) return;
    const double x0 = epi[0].pos[0], y0 = epi[0].pos[1], z0 = epi[0].pos[2];
    for (int i = 0; i < ni; i++) {
        const float xi = (float)(epi[i].pos[0] - x0);
        const float yi = (float)(epi[i].pos[1] - y0);
        const float zi = (float)(epi[i].pos[2] - z0);
        float ax = 0.f, ay = 0.f, az = 0.f, pot = 0.f;
        for (int j = 0; j < nj; j++) {
            const float rx = (float)(spj[j].pos[0] - x0) - xi;
            const float ry = (float)(spj[j].pos[1] - y0) - yi;
            const float rz = (float)(spj[j].pos[2] - z0) - zi;
            const float r2 = fmaf(rz, rz, fmaf(ry, ry, fmaf(rx, rx, eps2)));
            float r_inv = 1.0f / sqrtf(r2);
            const float tmp = 3.0f - r2 * (r_inv * r_inv);
            r_inv *= (tmp * 0.5f);
            const float m_r = (float)spj[j].mass * r_inv;
            const float m_r3 = m_r * r_inv * r_inv;
            ax += m_r3 * rx;
            ay += m_r3 * ry;
            az += m_r3 * rz;
            pot -= m_r;
        }
        force[i].acc[0] += ax;
        force[i].acc[1] += ay;
        force[i].acc[2] += az;
        force[i].phi += pot;
    }
}

/* ---------------------------------------------------------- FP64 twins */

/* Same algebra as orc_epep_f32 in double throughout (positions relative to the
 * group's first particle), neighbour predicate with the same 1.0201 factor. */
void orc_epep_f64(const gplum_epi *epi, int ni, const gplum_epj *epj, int nj,
                  double eps2, double *acc, double *phi, gplum_force *force)
{
    if (ni <= 0) return;
    const double x0 = epi[0].pos[0], y0 = epi[0].pos[1], z0 = epi[0].pos[2];
    for (int i = 0; i < ni; i++) {
        const double xi = epi[i].pos[0] - x0, yi = epi[i].pos[1] - y0, zi = epi[i].pos[2] - z0;
        const double roi = epi[i].r_out, rsi = epi[i].r_search;
        double ax = 0., ay = 0., az = 0., pot = 0.;
        int ngb = 0, rankf = 0, idmax = -1, idmin = 2147483647;
        for (int j = 0; j < nj; j++) {
            const double rx = (epj[j].pos[0] - x0) - xi;
            const double ry = (epj[j].pos[1] - y0) - yi;
            const double rz = (epj[j].pos[2] - z0) - zi;
            const double r2_real = rx * rx + ry * ry + rz * rz + eps2;
            const double rout = smax(roi, epj[j].r_out);
            const double rs = smax(rsi, epj[j].r_search);
            const double r2 = smax(r2_real, rout * rout);
            if (r2_real < rs * rs * 1.0201) {
                if (epi[i].id_local != epj[j].id_local || epi[i].myrank != epj[j].myrank) {
                    int rsub = epi[i].myrank - epj[j].myrank;
                    ngb++;
                    rankf += rsub * rsub;
                    if (epj[j].id_local > idmax) idmax = epj[j].id_local;
                    if (epj[j].id_local < idmin) idmin = epj[j].id_local;
                }
            }
            const double r_inv = 1.0 / sqrt(r2);
            const double mr_inv = epj[j].mass * r_inv;
            const double mr3_inv = r_inv * r_inv * mr_inv;
            ax += mr3_inv * rx;
            ay += mr3_inv * ry;
            az += mr3_inv * rz;
            pot -= mr_inv;
        }
        acc[3 * i + 0] += ax;
        acc[3 * i + 1] += ay;
        acc[3 * i + 2] += az;
        phi[i] += pot;
        if (force) {
            force[i].neighbor.number += ngb;
            force[i].neighbor.rank += rankf;
            if (idmax > force[i].neighbor.id_max) force[i].neighbor.id_max = idmax;
            if (idmin < force[i].neighbor.id_min) force[i].neighbor.id_min = idmin;
        }
    }
}

void orc_epsp_quad_f64(const gplum_epi *epi, int ni, const gplum_spj_quad *spj, int nj,
                       double eps2, double *acc, double *phi)
{
    if (ni <= 0) return;
    const double x0 = epi[0].pos[0], y0 = epi[0].pos[1], z0 = epi[0].pos[2];
    for (int i = 0; i < ni; i++) {
        const double xi = epi[i].pos[0] - x0, yi = epi[i].pos[1] - y0, zi = epi[i].pos[2] - z0;
        double ax = 0., ay = 0., az = 0., pot = 0.;
        for (int j = 0; j < nj; j++) {
            const double rx = (spj[j].pos[0] - x0) - xi;
            const double ry = (spj[j].pos[1] - y0) - yi;
            const double rz = (spj[j].pos[2] - z0) - zi;
            const double r2 = rx * rx + ry * ry + rz * rz + eps2;
            const double r_inv = 1.0 / sqrt(r2);
            const double r2_inv = r_inv * r_inv, r3_inv = r2_inv * r_inv;
            const double r4_inv = r2_inv * r2_inv, r5_inv = r2_inv * r3_inv;
            const double *q = spj[j].quad;
            const double tr = q[0] + q[1] + q[2];
            const double qxx = 3. * q[0] - tr, qyy = 3. * q[1] - tr, qzz = 3. * q[2] - tr;
            const double qxy = 3. * q[3], qzx = 3. * q[4], qyz = 3. * q[5];
            const double mtr = -(eps2 * tr);
            const double qrx = qxx * rx + qxy * ry + qzx * rz;
            const double qry = qyy * ry + qyz * rz + qxy * rx;
            const double qrz = qzz * rz + qzx * rx + qyz * ry;
            const double rqr = mtr + qrx * rx + qry * ry + qrz * rz;
            const double rqr_r4_inv = rqr * r4_inv;
            const double meff = spj[j].mass + 0.5 * rqr_r4_inv;
            const double meff3 = (spj[j].mass + 2.5 * rqr_r4_inv) * r3_inv;
            pot -= meff * r_inv;
            ax += -r5_inv * qrx + meff3 * rx;
            ay += -r5_inv * qry + meff3 * ry;
            az += -r5_inv * qrz + meff3 * rz;
        }
        acc[3 * i + 0] += ax;
        acc[3 * i + 1] += ay;
        acc[3 * i + 2] += az;
        phi[i] += pot;
    }
}

void orc_epsp_mono_f64(const gplum_epi *epi, int ni, const gplum_spj_mono *spj, int nj,
                       double eps2, double *acc, double *phi)
{
    if (ni <= 0) return;
    const double x0 = epi[0].pos[0], y0 = epi[0].pos[1], z0 = epi[0].pos[2];
    for (int i = 0; i < ni; i++) {
        const double xi = epi[i].pos[0] - x0, yi = epi[i].pos[1] - y0, zi = epi[i].pos[2] - z0;
        double ax = 0., ay = 0., az = 0., pot = 0.;
        for (int j = 0; j < nj; j++) {
            const double rx = (spj[j].pos[0] - x0) - xi;
            const double ry = (spj[j].pos[1] - y0) - yi;
            const double rz = (spj[j].pos[2] - z0) - zi;
            const double r2 = rx * rx + ry * ry + rz * rz + eps2;
            const double r_inv = 1.0 / sqrt(r2);
            const double m_r = spj[j].mass * r_inv;
            const double m_r3 = m_r * r_inv * r_inv;
            ax += m_r3 * rx;
            ay += m_r3 * ry;
            az += m_r3 * rz;
            pot -= m_r;
        }
        acc[3 * i + 0] += ax;
        acc[3 * i + 1] += ay;
        acc[3 * i + 2] += az;
        phi[i] += pot;
    }
}

/* ------------------------------------------- FP64 correction / decision */

/*
 * correctForceBetween2Particles (src/gravity_soft.h:75-153) and its Initial
 * variant (:155-242), with USE_RE_SEARCH_NEIGHBOR (main_p3t.cpp:15) and
 * USE_INDIVIDUAL_CUTOFF.  Operation order follows the reference; the FDPS vector
 * dot product is x*x'+y*y'+z*z' left to right.  t_min (units of length) uses
 * std::min/std::max, so dv2 == 0 gives NaN and dr2_min falls back to dr2.
 * The self pair (j_id == i_id) only adds m_j * r_out_inv to phi (:104-107).
 */
int orc_correct_pair(const gplum_params *p, const gplum_particle *pi,
                     const gplum_particle *pj, int initial, orc_corr_acc *a)
{
    const double eps2 = p->eps2;
    const double massj = pj->mass;
    const double r_out = smax(pi->r_out, pj->r_out);
    const double r_out_inv = smin(pi->r_out_inv, 1. / pj->r_out);
    const double r_search = smax(pi->r_search, pj->r_search);
    int accepted = 0;

    if (pj->id == pi->id) {
        a->phi += massj * r_out_inv;
        return 0;
    }
    const double drx = pj->pos[0] - pi->pos[0];
    const double dry = pj->pos[1] - pi->pos[1];
    const double drz = pj->pos[2] - pi->pos[2];
    double dr2 = drx * drx + dry * dry + drz * drz;
    dr2 += eps2;
    const double rij = sqrt(dr2);

    const double dvx = pj->vel[0] - pi->vel[0];
    const double dvy = pj->vel[1] - pi->vel[1];
    const double dvz = pj->vel[2] - pi->vel[2];
    const double dax = pj->acc_d[0] - pi->acc_d[0];
    const double day = pj->acc_d[1] - pi->acc_d[1];
    const double daz = pj->acc_d[2] - pi->acc_d[2];
    const double drdv = drx * dvx + dry * dvy + drz * dvz;
    const double dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
    const double da2 = dax * dax + day * day + daz * daz;
    const double t_min = smin(smax(-drdv / sqrt(dv2), 0.), p->dt_tree);
    const double dr2_min = smin(dr2, dr2 + 2. * drdv * t_min + dv2 * t_min * t_min);
    const double r_crit = p->R_search2 * r_out;
    const double v_crit_a = p->R_search3 * 0.5 * p->dt_tree;
    int cond = (dr2_min < r_crit * r_crit) || (dv2 < v_crit_a * v_crit_a * da2);
    if (initial) cond = cond || (da2 == 0.);
    if (cond) {
        if (rij < r_search) {
            a->number += 1;
            if (pj->id < a->id_cluster) a->id_cluster = pj->id;
            a->acc0 += r_out * r_out / massj;
            accepted = 1;
        }
    }
    if (rij < r_out) {
        const double rinv = 1. / rij;
        const double r2inv = rinv * rinv;
        const double r3inv = rinv * r2inv;
        const double W = orc_cutoff_W(p, rij, r_out_inv);
        const double K = orc_cutoff_K(p, rij, r_out_inv);
        const double r_min = smin(rinv, r_out_inv);
        a->phi -= massj * (rinv * W - r_min);
        const double cf = massj * (r3inv * K - r_min * r_min * r_min);
        a->acc[0] += cf * drx;
        a->acc[1] += cf * dry;
        a->acc[2] += cf * drz;
        if (initial) {
            const double alpha = drdv * r2inv;
            const double dKdt = orc_cutoff_dKdt(p, rij, r_out_inv, alpha);
            const double alpha_c = alpha * (1. - K);
            a->phi_d -= massj * rinv * (1. - W);
            const double cd = massj * r3inv * (1. - K);
            a->acc_d[0] += cd * drx;
            a->acc_d[1] += cd * dry;
            a->acc_d[2] += cd * drz;
            const double mr3 = massj * r3inv;
            const double c2 = 3. * alpha_c + dKdt;
            a->jerk[0] += mr3 * ((1. - K) * dvx - c2 * drx);
            a->jerk[1] += mr3 * ((1. - K) * dvy - c2 * dry);
            a->jerk[2] += mr3 * ((1. - K) * dvz - c2 * drz);
        }
    }
    return accepted;
}

void orc_nlist_free(orc_nlist *nl)
{
    if (!nl) return;
    free(nl->offset);
    free(nl->index);
    nl->offset = NULL;
    nl->index = NULL;
    nl->n = 0;
}
