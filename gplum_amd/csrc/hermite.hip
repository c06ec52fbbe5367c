// hermite.hip -- batched FP64 4th-order Hermite with block time steps (hard part),
// Kepler drift and Sun-only Hermite for isolated particles.
//
// One workgroup integrates one cluster over local time [0, dt_tree] exactly as
// timeIntegrate_multi does (src/hermite.h:224-479): a block step picks the minimum
// time+dt, predicts every particle, checks collisions (perfect merging), corrects
// the active particles with two P(EC) iterations and updates their block steps.
// The active particles of one block step are independent (each reads only the
// predicted positions of the others), so they are corrected in parallel with one
// barrier per block step.  The j-data of the cluster (predicted position/velocity,
// mass, 1/r_out, id) lives in LDS for clusters of up to HERMITE_LDS particles.
// The kernel is persistent: workgroups pull clusters (costliest first: estimated block
// steps times size) from an atomic queue until it is empty.
#include <algorithm>
#include <cfloat>
#include <mutex>
#include <vector>

#include "hermite.hpp"

namespace gplum {

#define SECONDORDER 5.e-1
#define THIRDORDER 1.6666666666666667e-1
#define ALPHA_POS 1.1666666666666667e-3  // ALPHA = 7/6000 of the corrector's position term, src/particle.h:339 (kept as written)
#define ONE_TWELFTH 8.3333333333333333e-2
#define ONE_SIXTIETH 1.6666666666666667e-2

struct HP {                                  // physics constants for the kernels
    double m_sun, eps2, eps2_sun, dt_tree, dt_min, eta, eta_0, eta_sun, eta_sun0, alpha2;
    Cutoff cut;
    int collision;
    int terms_max;     // clusters of 2 .. terms_max (<= 4) bodies run integrate_terms
};

__device__ inline double dmin(double a, double b) { return (b < a) ? b : a; }

// collisionDetermination looks for R = |dx| / (f_i R_i + f_j R_j) < 1.  Pairs with
// |dx|^2 >= (1 + 1e-10) (f_i R_i + f_j R_j)^2 have R >= 1 even after the rounding of
// sqrt and the division, so they are skipped before computing R; the minimum and its
// pair are unchanged.
constexpr double COL_FILTER = 1.0 + 1e-10;

// ---------------------------------------------------------- Sun gravity

__device__ inline void star_acc_jerk(const HP &P, const double x[3], const double v[3],
                                     double acc[3], double jerk[3], double *phi) {
    const double dr[3] = {-x[0], -x[1], -x[2]};
    const double dv[3] = {-v[0], -v[1], -v[2]};
    const double r2inv = 1. / (dot3(dr[0], dr[1], dr[2], dr[0], dr[1], dr[2]) + P.eps2_sun);
    const double rinv = sqrt(r2inv);
    const double r3inv = rinv * r2inv;
    const double mj = P.m_sun * r3inv;
    const double alpha = dot3(dr[0], dr[1], dr[2], dv[0], dv[1], dv[2]) * r2inv;
    if (phi) *phi = -P.m_sun * rinv;
    for (int k = 0; k < 3; k++) {
        acc[k] = mj * dr[k];
        jerk[k] = mj * (dv[k] - 3. * alpha * dr[k]);
    }
}

// -------------------------------------------------------------- time steps

__device__ inline double calc_dt2(double eta, double alpha2, double acc0, const double a[3],
                                  const double j[3]) {
    const double Acc2 = dot3(a[0], a[1], a[2], a[0], a[1], a[2]) + alpha2 * acc0 * acc0;
    const double Jerk2 = dot3(j[0], j[1], j[2], j[0], j[1], j[2]);
    return (Jerk2 > 0.) ? eta * sqrt(Acc2 / Jerk2) : DBL_MAX;
}
__device__ inline double calc_dt4(double eta, double alpha2, double acc0, const double a[3],
                                  const double j[3], const double s[3], const double c[3]) {
    const double Acc = sqrt(dot3(a[0], a[1], a[2], a[0], a[1], a[2]) + alpha2 * acc0 * acc0);
    const double Jerk2 = dot3(j[0], j[1], j[2], j[0], j[1], j[2]);
    const double Jerk = sqrt(Jerk2);
    const double Snap2 = dot3(s[0], s[1], s[2], s[0], s[1], s[2]);
    const double Snap = sqrt(Snap2);
    const double Crac = sqrt(dot3(c[0], c[1], c[2], c[0], c[1], c[2]));
    return (Jerk > 0.) ? eta * sqrt((Acc * Snap + Jerk2) / (Jerk * Crac + Snap2)) : DBL_MAX;
}

// FPGrav::calcDeltatInitial, src/particle.h:886-914.
__device__ inline double dt_initial(const HP &P, double time, double dt, double acc0,
                                    const double acc_d[3], const double jerk_d[3],
                                    const double acc_s[3], const double jerk_s[3]) {
    double dt_next = 0.5 * P.dt_tree;
    const double dt_1 = dmin(calc_dt2(P.eta_0, P.alpha2, acc0, acc_d, jerk_d),
                             calc_dt2(P.eta_sun0, P.alpha2, 0., acc_s, jerk_s));
    double rem = fmod_block(time, dt_next);
    for (int it = 0; rem != 0.0 && it < HALVE_CAP; it++) { dt_next *= 0.5; rem = fmod_block(time, dt_next); }
    if (rem != 0.0) return P.dt_min;
    if (dt > 0.)
        for (int it = 0; 2. * dt < dt_next && it < HALVE_CAP; it++) dt_next *= 0.5;
    for (int it = 0; dt_1 < dt_next && it < HALVE_CAP; it++) dt_next *= 0.5;
    if (dt_next < 2. * P.dt_min) dt_next = P.dt_min;
    return dt_next;
}

// calc_dt4 with the hardware square root and reciprocal (v_sqrt_f64, v_rcp_f64, not
// correctly rounded; Acc * Snap as sqrt(Acc2 * Snap2), Jerk * Crac likewise): within a
// few 1e-7 of calc_dt4 relative, DBL_MAX exactly where calc_dt4 gives it (Jerk > 0 iff
// Jerk2 > 0).  A product that underflows (0 from non-zero factors, or subnormal) gives
// NaN and one that overflows gives inf or NaN; dt_block sends either to calc_dt4.
__device__ inline double calc_dt4_fast(double eta, double alpha2, double acc0, const double a[3],
                                       const double j[3], const double s[3], const double c[3]) {
    const double Acc2 = dot3(a[0], a[1], a[2], a[0], a[1], a[2]) + alpha2 * acc0 * acc0;
    const double Jerk2 = dot3(j[0], j[1], j[2], j[0], j[1], j[2]);
    const double Snap2 = dot3(s[0], s[1], s[2], s[0], s[1], s[2]);
    const double Crac2 = dot3(c[0], c[1], c[2], c[0], c[1], c[2]);
    // selects, not branches: the Sun and planet criteria of dt_block stay one basic block
    const double as2 = Acc2 * Snap2, jc2 = Jerk2 * Crac2;
    const bool ok = (as2 >= DBL_MIN || (Acc2 == 0. || Snap2 == 0.)) && (jc2 >= DBL_MIN || Crac2 == 0.);
    const double num = __builtin_amdgcn_sqrt(as2) + Jerk2;
    const double den = __builtin_amdgcn_sqrt(jc2) + Snap2;
    const double q = ok ? eta * __builtin_amdgcn_sqrt(num * __builtin_amdgcn_rcp(den)) : __builtin_nan("");
    return (Jerk2 > 0.) ? q : DBL_MAX;
}

// True when t is farther than 2^-10 relative from every d * 2^k (k integer), for
// t > 2^-1000 and finite: then t and any value within 2^-11 of it relative lie on the
// same side of each of those grid points.  t = m_t 2^e_t, d = m_d 2^e_d with mantissas
// in [0.5, 1): the grid points nearest t are at m_t / m_d in {1/2, 1, 2}.
__device__ inline bool clear_of_grid(double t, double d) {
    int et = 0, ed = 0;
    const double r = frexp(t, &et) * __builtin_amdgcn_rcp(frexp(d, &ed));
    const double tol = 0x1p-10;
    return (t > 0x1p-1000 && t < DBL_MAX) && r > 0.5 * (1. + tol) && (r < 1. - tol || r > 1. + tol) &&
           r < 2. * (1. - tol);
}

// FPHard::calcDeltat, src/particle.h:1276-1314.  dt_1 enters the result only through
// dt_1 < dt_next for dt_next = D, D/2, D/4, ... (D the step left by the fmod loop,
// halvings exact above 2^-1000), so a dt_1 from calc_dt4_fast that is clear of that
// grid decides every comparison as the correctly rounded dt_1 does; near the grid, or
// when either fast criterion is NaN or infinite (under- or overflow of its products,
// which the minimum would otherwise drop), calc_dt4 itself is evaluated.  Same result,
// bit for bit.
// FAST = false: calc_dt4 always (the self-check's reference).
template <bool FAST = true>
__device__ inline double dt_block(const HP &P, double time, double dt, double acc0,
                                  const double acc_d[3], const double jerk_d[3], const double a2_d[3],
                                  const double a3_d[3], const double acc_s[3], const double jerk_s[3],
                                  const double a2_s[3], const double a3_s[3], int *fell_back = nullptr) {
    double dt_next = dmin(2. * dt, 0.5 * P.dt_tree);
    double rem = fmod_block(time, dt_next);
    for (int it = 0; rem != 0.0 && it < HALVE_CAP; it++) { dt_next *= 0.5; rem = fmod_block(time, dt_next); }
    if (rem != 0.0) return P.dt_min;
    double dt_1 = 0.;
    bool exact = !FAST;
    if (FAST) {
        const double f_planet = calc_dt4_fast(P.eta, P.alpha2, acc0, acc_d, jerk_d, a2_d, a3_d);
        const double f_sun = calc_dt4_fast(P.eta_sun, P.alpha2, 0., acc_s, jerk_s, a2_s, a3_s);
        dt_1 = dmin(f_planet, f_sun);
        // DBL_MAX is a valid value (no jerk); NaN and inf are not
        exact = !(f_planet <= DBL_MAX && f_sun <= DBL_MAX) || !clear_of_grid(dt_1, dt_next);
        if (fell_back) *fell_back = exact ? 1 : 0;
    }
    if (exact)
        dt_1 = dmin(calc_dt4(P.eta, P.alpha2, acc0, acc_d, jerk_d, a2_d, a3_d),
                    calc_dt4(P.eta_sun, P.alpha2, 0., acc_s, jerk_s, a2_s, a3_s));
    for (int it = 0; dt_1 < dt_next && 0.5 * dt < dt_next && it < HALVE_CAP; it++) dt_next *= 0.5;
    if (dt_next < 2. * P.dt_min) dt_next = P.dt_min;
    return dt_next;
}

// ------------------------------------------------------------- hard pair

// gravity_hard.h:219-255: the summands of acc_d, jerk_d (and phi_d) of i from j,
// returned separately so that several j can be evaluated side by side and added in
// ascending j afterwards (the sums stay those of a sequential loop).
__device__ inline void pair_term(const HP &P, const double xi[3], const double vi[3], double rinv_i,
                                 double xj0, double xj1, double xj2, double vj0, double vj1, double vj2,
                                 double mj, double rinv_j, double ta[3], double tj[3], double *tphi) {
    const double dr[3] = {xj0 - xi[0], xj1 - xi[1], xj2 - xi[2]};
    double dr2 = dot3(dr[0], dr[1], dr[2], dr[0], dr[1], dr[2]);
    dr2 += P.eps2;
    const double dv[3] = {vj0 - vi[0], vj1 - vi[1], vj2 - vi[2]};
    const double rij = sqrt(dr2);
    const double rinv = 1. / rij;
    const double r2inv = rinv * rinv;
    const double r3inv = r2inv * rinv;
    const double r_out_inv = dmin(rinv_i, rinv_j);
    const double mj_rij3 = mj * r3inv;
    const double alpha = dot3(dr[0], dr[1], dr[2], dv[0], dv[1], dv[2]) * r2inv;
    const double _K = 1. - cutoff_K(P.cut, rij, r_out_inv);
    const double dKdt = cutoff_dKdt(P.cut, rij, r_out_inv, alpha);
    const double alpha_c = alpha * _K;
    if (tphi) {
        const double _W = 1. - cutoff_W(P.cut, rij, r_out_inv);
        *tphi = mj * rinv * _W;
    }
    const double c2 = 3. * alpha_c + dKdt;
    for (int k = 0; k < 3; k++) {
        ta[k] = mj_rij3 * _K * dr[k];
        tj[k] = mj_rij3 * (_K * dv[k] - c2 * dr[k]);
    }
}

// ---------------------------------------------------------- cluster state

struct HardView {          // SoA arrays of the gathered cluster particles
    double *pos, *vel, *acc_s, *jerk_s, *acc_d, *jerk_d, *xp, *vp, *a0_s, *j0_s, *a0_d, *j0_d;
    double *a2_s, *a3_s, *a2_d, *a3_d;                       // 3 per particle
    double *mass, *rinv, *time, *dt, *acc0, *r_planet, *f, *phi_d, *phi_s;
    int64_t *id;
    int32_t *flags, *twin_of, *act;     // act: active in the current block step
    // collision records (COLREC doubles each, see perfect_merge), when the caller keeps a
    // collision log: crec null otherwise
    double *crec = nullptr;
    unsigned *ccount = nullptr;
    unsigned ccap = 0;
};

__device__ inline void ld3(const double *a, int i, double v[3]) {
    v[0] = a[3 * i]; v[1] = a[3 * i + 1]; v[2] = a[3 * i + 2];
}
__device__ inline void st3(double *a, int i, const double v[3]) {
    a[3 * i] = v[0]; a[3 * i + 1] = v[1]; a[3 * i + 2] = v[2];
}

// j-data accessor: LDS when the cluster fits, else global memory.
struct JData {
    const double *x, *v, *m, *ri;
    const int64_t *id;
    int stride;            // 3 for interleaved global arrays
};

// pairs evaluated side by side per lane (independent dependency chains; the one wave
// of a small cluster has no other wave to hide FP64 latency behind)
constexpr int HARD_ILP = 4;

__device__ inline void pair_term_j(const HP &P, const JData &J, int j, const double xi[3], const double vi[3],
                                   double rinv_i, double ta[3], double tj[3], double *tphi) {
    pair_term(P, xi, vi, rinv_i, J.x[3 * j], J.x[3 * j + 1], J.x[3 * j + 2], J.v[3 * j], J.v[3 * j + 1],
              J.v[3 * j + 2], J.m[j], J.ri[j], ta, tj, tphi);
}

// Sum over j = j0, j0 + s, j0 + 2s, ... < n (j != i by id) in ascending j, ILP pairs
// evaluated at a time.  The self pair is evaluated too (its summands may be NaN) and
// not added.
template <int ILP = HARD_ILP>
__device__ inline void pair_sum_strided(const HP &P, const JData &J, int j0, int s, int n, int64_t idi,
                                        const double xi[3], const double vi[3], double rinv_i, double acc[3],
                                        double jerk[3], double *phi) {
    int j = j0;
    for (; j + (ILP - 1) * s < n; j += ILP * s) {
        double ta[ILP][3], tj[ILP][3], tp[ILP];
#pragma unroll
        for (int u = 0; u < ILP; u++) pair_term_j(P, J, j + u * s, xi, vi, rinv_i, ta[u], tj[u], phi ? &tp[u] : nullptr);
#pragma unroll
        for (int u = 0; u < ILP; u++) {
            if (J.id[j + u * s] == idi) continue;
            for (int k = 0; k < 3; k++) { acc[k] += ta[u][k]; jerk[k] += tj[u][k]; }
            if (phi) *phi -= tp[u];
        }
    }
    for (; j < n; j += s) {
        if (J.id[j] == idi) continue;
        double ta[3], tj[3], tp;
        pair_term_j(P, J, j, xi, vi, rinv_i, ta, tj, phi ? &tp : nullptr);
        for (int k = 0; k < 3; k++) { acc[k] += ta[k]; jerk[k] += tj[k]; }
        if (phi) *phi -= tp;
    }
}

__device__ inline void gravity_loop(const HP &P, const JData &J, int n, int64_t idi,
                                    const double xi[3], const double vi[3], double rinv_i,
                                    double acc[3], double jerk[3], double *phi) {
    acc[0] = acc[1] = acc[2] = 0.;
    jerk[0] = jerk[1] = jerk[2] = 0.;
    pair_sum_strided(P, J, 0, 1, n, idi, xi, vi, rinv_i, acc, jerk, phi);
}

// gravity_loop for the one-particle-per-lane path of lane i: the lane's partners are
// j = k + (k >= i) for k = 0 .. n-2, ascending j as in gravity_loop, so every lane runs
// the same n-1 iterations (the loop over all j with a self test would run n iterations
// with one lane masked off in each).  A partner with the same id (never, before any
// merger) is evaluated and not added, as gravity_loop skips it; CHECK = false leaves
// the test out, for clusters whose ids are known to be distinct.
template <bool CHECK = true>
__device__ inline void gravity_loop_lane(const HP &P, const JData &J, int n, int i, int64_t idi,
                                         const double xi[3], const double vi[3], double rinv_i,
                                         double acc[3], double jerk[3], double *phi) {
    acc[0] = acc[1] = acc[2] = 0.;
    jerk[0] = jerk[1] = jerk[2] = 0.;
    const int m = n - 1;
    int k = 0;
    for (; k + HARD_ILP <= m; k += HARD_ILP) {
        double ta[HARD_ILP][3], tj[HARD_ILP][3], tp[HARD_ILP];
        int jj[HARD_ILP];
#pragma unroll
        for (int u = 0; u < HARD_ILP; u++) {
            jj[u] = k + u + ((k + u) >= i ? 1 : 0);
            pair_term_j(P, J, jj[u], xi, vi, rinv_i, ta[u], tj[u], phi ? &tp[u] : nullptr);
        }
#pragma unroll
        for (int u = 0; u < HARD_ILP; u++) {
            const bool use = !CHECK || J.id[jj[u]] != idi;
            for (int d = 0; d < 3; d++) {
                acc[d] = use ? acc[d] + ta[u][d] : acc[d];
                jerk[d] = use ? jerk[d] + tj[u][d] : jerk[d];
            }
            if (phi) *phi = use ? *phi - tp[u] : *phi;
        }
    }
    for (; k < m; k++) {
        const int j = k + (k >= i ? 1 : 0);
        double ta[3], tj[3], tp;
        pair_term_j(P, J, j, xi, vi, rinv_i, ta, tj, phi ? &tp : nullptr);
        const bool use = !CHECK || J.id[j] != idi;
        for (int d = 0; d < 3; d++) {
            acc[d] = use ? acc[d] + ta[d] : acc[d];
            jerk[d] = use ? jerk[d] + tj[d] : jerk[d];
        }
        if (phi) *phi = use ? *phi - tp : *phi;
    }
}

// Partner j's predicted state (position, velocity, mass, 1/r_out), held in registers
// for the three force evaluations of a block step.
struct JReg {
    double x[3], v[3], m, ri;
};
__device__ inline void load_jreg(const JData &J, int j, JReg &r) {
    for (int d = 0; d < 3; d++) { r.x[d] = J.x[3 * j + d]; r.v[d] = J.v[3 * j + d]; }
    r.m = J.m[j];
    r.ri = J.ri[j];
}

// force_lane_fixed<M> on partner data in registers.
template <int M>
__device__ inline void force_jreg(const HP &P, const JReg (&pj)[M], const double xi[3], const double vi[3],
                                  double rinv_i, double as[3], double js[3], double ad[3], double jd[3]) {
    double ta[M][3], tj[M][3];
#pragma unroll
    for (int k = 0; k < M; k++)
        pair_term(P, xi, vi, rinv_i, pj[k].x[0], pj[k].x[1], pj[k].x[2], pj[k].v[0], pj[k].v[1], pj[k].v[2],
                  pj[k].m, pj[k].ri, ta[k], tj[k], nullptr);
    star_acc_jerk(P, xi, vi, as, js, nullptr);
    for (int d = 0; d < 3; d++) { ad[d] = 0.; jd[d] = 0.; }
#pragma unroll
    for (int k = 0; k < M; k++)
        for (int d = 0; d < 3; d++) { ad[d] += ta[k][d]; jd[d] += tj[k][d]; }
}

// calcGravity (full, with phi_d and Sun phi), gravity_hard.h:203-256, current pos/vel.
__device__ void gravity_full(const HP &P, HardView &H, int b, int n, int i) {
    double x[3], v[3], as[3], js[3], ad[3], jd[3], phs, phd = 0.;
    ld3(H.pos, b + i, x);
    ld3(H.vel, b + i, v);
    star_acc_jerk(P, x, v, as, js, &phs);
    const JData Jc{H.pos + 3 * b, H.vel + 3 * b, H.mass + b, H.rinv + b, H.id + b, 3};
    gravity_loop(P, Jc, n, H.id[b + i], x, v, H.rinv[b + i], ad, jd, &phd);
    st3(H.acc_s, b + i, as);
    st3(H.jerk_s, b + i, js);
    st3(H.acc_d, b + i, ad);
    st3(H.jerk_d, b + i, jd);
    H.phi_s[b + i] = phs;
    H.phi_d[b + i] = phd;
}

// FPHard::predict, src/particle.h:1183-1207.
__device__ inline void predict(HardView &H, int g, double Dt) {
    for (int k = 0; k < 3; k++) {
        const double pos = H.pos[3 * g + k], vel = H.vel[3 * g + k];
        const double as = H.acc_s[3 * g + k], ad = H.acc_d[3 * g + k];
        const double js = H.jerk_s[3 * g + k], jd = H.jerk_d[3 * g + k];
        H.a0_s[3 * g + k] = as;
        H.j0_s[3 * g + k] = js;
        H.j0_d[3 * g + k] = jd;
        H.a0_d[3 * g + k] = ad;
        const double acc_sd = as + ad;
        const double jerk_sd = js + jd;
        H.xp[3 * g + k] = pos + Dt * (vel + Dt * (SECONDORDER * acc_sd + THIRDORDER * Dt * jerk_sd));
        H.vp[3 * g + k] = vel + Dt * (acc_sd + Dt * SECONDORDER * jerk_sd);
    }
}

// FPHard::correct, src/particle.h:1208-1274 (position term with ALPHA = 7/6000, as written).
__device__ inline void correct(HardView &H, int g, double Dt) {
    const double Dt2 = Dt * Dt;
    const double Dt_inv = recip_block(Dt);
    const double Dt_inv2 = Dt_inv * Dt_inv;
    const double Dt_inv3 = Dt_inv2 * Dt_inv;
    for (int k = 0; k < 3; k++) {
        const int q = 3 * g + k;
        const double Am_s = H.acc_s[q] - H.a0_s[q];
        const double J0_s = Dt * H.j0_s[q];
        const double J1_s = Dt * H.jerk_s[q];
        const double A2_s = 3. * Am_s - (J1_s + 2. * J0_s);
        const double A3_s = -2. * Am_s + (J1_s + J0_s);
        H.a2_s[q] = 2. * Dt_inv2 * (A2_s + 3. * A3_s);
        H.a3_s[q] = 6. * Dt_inv3 * A3_s;
        const double Am_d = H.acc_d[q] - H.a0_d[q];
        const double J0_d = H.j0_d[q] * Dt;
        const double J1_d = H.jerk_d[q] * Dt;
        const double A2_d = 3. * Am_d - (J1_d + 2. * J0_d);
        const double A3_d = -2. * Am_d + (J1_d + J0_d);
        H.a2_d[q] = 2. * (A2_d + 3. * A3_d) * Dt_inv2;
        H.a3_d[q] = 6. * A3_d * Dt_inv3;
        const double A2_sd = A2_s + A2_d;
        const double A3_sd = A3_s + A3_d;
        H.pos[q] = H.xp[q] + ONE_SIXTIETH * Dt2 * (5. * A2_sd + 3. * ALPHA_POS * A3_sd);
        H.vel[q] = H.vp[q] + ONE_TWELFTH * Dt * (4. * A2_sd + 3. * A3_sd);
    }
}

// Register-resident twins of predict() and correct() (same operation order), for the
// one-particle-per-lane path of clusters that fit one wave.
__device__ inline void predict_regs(double Dt, const double pos[3], const double vel[3], const double as[3],
                                    const double js[3], const double ad[3], const double jd[3], double a0s[3],
                                    double j0s[3], double a0d[3], double j0d[3], double xp[3], double vp[3]) {
    for (int k = 0; k < 3; k++) {
        a0s[k] = as[k];
        j0s[k] = js[k];
        j0d[k] = jd[k];
        a0d[k] = ad[k];
        const double acc_sd = as[k] + ad[k];
        const double jerk_sd = js[k] + jd[k];
        xp[k] = pos[k] + Dt * (vel[k] + Dt * (SECONDORDER * acc_sd + THIRDORDER * Dt * jerk_sd));
        vp[k] = vel[k] + Dt * (acc_sd + Dt * SECONDORDER * jerk_sd);
    }
}

__device__ inline void correct_regs(double Dt, const double xp[3], const double vp[3], const double as[3],
                                    const double a0s[3], const double j0s[3], const double js[3], const double ad[3],
                                    const double a0d[3], const double j0d[3], const double jd[3], double pos[3],
                                    double vel[3], double a2s[3], double a3s[3], double a2d[3], double a3d[3]) {
    const double Dt2 = Dt * Dt;
    const double Dt_inv = recip_block(Dt);
    const double Dt_inv2 = Dt_inv * Dt_inv;
    const double Dt_inv3 = Dt_inv2 * Dt_inv;
    for (int k = 0; k < 3; k++) {
        const double Am_s = as[k] - a0s[k];
        const double J0_s = Dt * j0s[k];
        const double J1_s = Dt * js[k];
        const double A2_s = 3. * Am_s - (J1_s + 2. * J0_s);
        const double A3_s = -2. * Am_s + (J1_s + J0_s);
        a2s[k] = 2. * Dt_inv2 * (A2_s + 3. * A3_s);
        a3s[k] = 6. * Dt_inv3 * A3_s;
        const double Am_d = ad[k] - a0d[k];
        const double J0_d = j0d[k] * Dt;
        const double J1_d = jd[k] * Dt;
        const double A2_d = 3. * Am_d - (J1_d + 2. * J0_d);
        const double A3_d = -2. * Am_d + (J1_d + J0_d);
        a2d[k] = 2. * (A2_d + 3. * A3_d) * Dt_inv2;
        a3d[k] = 6. * A3_d * Dt_inv3;
        const double A2_sd = A2_s + A2_d;
        const double A3_sd = A3_s + A3_d;
        pos[k] = xp[k] + ONE_SIXTIETH * Dt2 * (5. * A2_sd + 3. * ALPHA_POS * A3_sd);
        vel[k] = vp[k] + ONE_TWELFTH * Dt * (4. * A2_sd + 3. * A3_sd);
    }
}

// gravity_p / gravity_c of particle g (cluster-local i) against the predicted j-data.
__device__ inline void gravity_pc(const HP &P, HardView &H, const JData &J, int n, int g,
                                  bool corrected) {
    double x[3], v[3], as[3], js[3], ad[3], jd[3];
    if (corrected) { ld3(H.pos, g, x); ld3(H.vel, g, v); }
    else { ld3(H.xp, g, x); ld3(H.vp, g, v); }
    star_acc_jerk(P, x, v, as, js, nullptr);
    gravity_loop(P, J, n, H.id[g], x, v, H.rinv[g], ad, jd, nullptr);
    st3(H.acc_s, g, as);
    st3(H.jerk_s, g, js);
    st3(H.acc_d, g, ad);
    st3(H.jerk_d, g, jd);
}

// Hard force on one particle with all 64 lanes of the (one-wave) workgroup: lane l
// sums the pairs j = l, l+64, ... in ascending j, then a butterfly over the lanes adds
// the 64 partial sums (every lane ends with the same total).  This fixed order is the
// hard-force summation order of clusters above HARD_SEQ_MAX particles (the oracle sums
// the same way); smaller clusters sum sequentially in j.
// Lane permutations within rows of 16 lanes (DPP): quad_perm [1,0,3,2], quad_perm
// [2,3,0,1], row_half_mirror, row_mirror.  Applied in this order with a min, they
// leave every lane of a row holding the row's minimum.
constexpr int DPP_QUAD_1032 = 0xB1, DPP_QUAD_2301 = 0x4E, DPP_ROW_HALF_MIRROR = 0x141, DPP_ROW_MIRROR = 0x140;
// quad_perm [3,2,1,0] and row_ror:8 (rotate a row of 16 lanes by 8: lane ^ 8)
constexpr int DPP_QUAD_3210 = 0x1B, DPP_ROW_ROR8 = 0x128;

template <int CTRL>
__device__ inline double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ inline double readlane_f64(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ inline void coop_pair_sum(const HP &P, const JData &J, int n, int64_t idi, const double xi[3],
                                     const double vi[3], double rinv_i, double acc[3], double jerk[3]) {
    double a[3] = {0., 0., 0.}, jk[3] = {0., 0., 0.};
    // two pairs at a time: the multi-wave kernels hold more state per lane
    pair_sum_strided<2>(P, J, (int)(threadIdx.x & 63), 64, n, idi, xi, vi, rinv_i, a, jk, nullptr);
    // butterfly x += x[lane ^ d] for d = 32, 16, 8, 4, 2, 1 (every lane active): the two
    // levels across rows of 16 lanes by shuffles, the four inside a row by DPP moves
    // (row_ror:8 is lane ^ 8; quad_perm [3,2,1,0] after row_half_mirror is lane ^ 4;
    // quad_perm [2,3,0,1] and [1,0,3,2] are lane ^ 2 and lane ^ 1), which skip the LDS
    // crossbar.  The same partner values in the same order as the shuffle butterfly.
    for (int d = 32; d > 8; d >>= 1)
        for (int k = 0; k < 3; k++) {
            a[k] += __shfl_xor(a[k], d);
            jk[k] += __shfl_xor(jk[k], d);
        }
    for (int k = 0; k < 3; k++) { a[k] += dpp_f64<DPP_ROW_ROR8>(a[k]); jk[k] += dpp_f64<DPP_ROW_ROR8>(jk[k]); }
    for (int k = 0; k < 3; k++) {
        a[k] += dpp_f64<DPP_QUAD_3210>(dpp_f64<DPP_ROW_HALF_MIRROR>(a[k]));
        jk[k] += dpp_f64<DPP_QUAD_3210>(dpp_f64<DPP_ROW_HALF_MIRROR>(jk[k]));
    }
    for (int k = 0; k < 3; k++) { a[k] += dpp_f64<DPP_QUAD_2301>(a[k]); jk[k] += dpp_f64<DPP_QUAD_2301>(jk[k]); }
    for (int k = 0; k < 3; k++) { a[k] += dpp_f64<DPP_QUAD_1032>(a[k]); jk[k] += dpp_f64<DPP_QUAD_1032>(jk[k]); }
    for (int k = 0; k < 3; k++) { acc[k] = a[k]; jerk[k] = jk[k]; }
}

// mergeAccJerk (hermite.h:36-74) for target t: the mass-weighted mean of the target's
// and its twins' acc_d/jerk_d (each computed with its own cut-off radius), in twin
// index order, written to all of them.
__device__ inline void merge_acc_jerk(HardView &H, int b, int n, int t) {
    double acc[3], jerk[3], mass = H.mass[b + t];
    for (int k = 0; k < 3; k++) {
        acc[k] = H.mass[b + t] * H.acc_d[3 * (b + t) + k];
        jerk[k] = H.mass[b + t] * H.jerk_d[3 * (b + t) + k];
    }
    for (int j = 0; j < n; j++) {
        if (H.twin_of[b + j] != t) continue;
        for (int k = 0; k < 3; k++) {
            acc[k] += H.mass[b + j] * H.acc_d[3 * (b + j) + k];
            jerk[k] += H.mass[b + j] * H.jerk_d[3 * (b + j) + k];
        }
        mass += H.mass[b + j];
    }
    const double minv = 1. / mass;
    for (int k = 0; k < 3; k++) { acc[k] *= minv; jerk[k] *= minv; }
    st3(H.acc_d, b + t, acc);
    st3(H.jerk_d, b + t, jerk);
    for (int j = 0; j < n; j++) {
        if (H.twin_of[b + j] != t) continue;
        st3(H.acc_d, b + j, acc);
        st3(H.jerk_d, b + j, jerk);
    }
}

// Perfect merging (collisionA.h:91-103, 155-356, 358-669; hermite.h:413-423), run
// by one thread.  Returns the dissipated energy.
__device__ double perfect_merge(const HP &P, HardView &H, int b, int n, int ic, int jc) {
    const int I = b + ic, T = b + jc;
    double pos_imp[3], pos_tar[3], vel_imp[3], vel_tar[3];
    ld3(H.pos, I, pos_imp); ld3(H.pos, T, pos_tar);
    ld3(H.vel, I, vel_imp); ld3(H.vel, T, vel_tar);
    double mass_imp = H.mass[I], mass_tar = H.mass[T];
    if (H.flags[I] & GPLUM_FLAG_MERGED)
        for (int a = 0; a < n; a++) if (H.twin_of[b + a] == ic) mass_imp += H.mass[b + a];
    if (H.flags[T] & GPLUM_FLAG_MERGED)
        for (int a = 0; a < n; a++) if (H.twin_of[b + a] == jc) mass_tar += H.mass[b + a];
    double vel_g[3], pos_new[3], vel_new[3];
    for (int k = 0; k < 3; k++) {
        vel_g[k] = (mass_imp * vel_imp[k] + mass_tar * vel_tar[k]) / (mass_imp + mass_tar);
        pos_new[k] = (mass_imp * pos_imp[k] + mass_tar * pos_tar[k]) / (mass_imp + mass_tar);
        vel_new[k] = vel_g[k];
    }
    if (H.crec) {
        // Collision0::write2File's record (collisionA.h:122-148; inputPair, setParticle
        // :155-356 with n_frag = 0): the impactor's cluster-local time, ids, masses with
        // their twins, positions and velocities at the collision, and the relative shifts
        // of the centre of mass, |g_new - g| / |g| for position and velocity, where g =
        // pos_g of inputPair (= pos_new here) and g_new = ((m_imp - 0) pos_new + m_tar
        // pos_new) / (m_imp - 0 + m_tar) of setParticle.  Slot COLREC - 1: the cluster's
        // base index, for the log's cluster order.
        const unsigned k = atomicAdd(H.ccount, 1u);
        if (k < H.ccap) {
            double *r = H.crec + (size_t)k * COLREC;
            double dp[3], dv[3];
            const double mt = mass_imp - 0. + mass_tar;
            for (int q = 0; q < 3; q++) {
                dp[q] = ((mass_imp - 0.) * pos_new[q] + mass_tar * pos_new[q]) / mt - pos_new[q];
                dv[q] = ((mass_imp - 0.) * vel_new[q] + mass_tar * vel_new[q]) / mt - vel_g[q];
            }
            r[0] = H.time[I];
            r[1] = (double)H.id[I];
            r[2] = (double)H.id[T];
            r[3] = mass_imp;
            r[4] = mass_tar;
            for (int q = 0; q < 3; q++) {
                r[5 + q] = pos_imp[q]; r[8 + q] = pos_tar[q];
                r[11 + q] = vel_imp[q]; r[14 + q] = vel_tar[q];
            }
            r[17] = sqrt(dot3(dp[0], dp[1], dp[2], dp[0], dp[1], dp[2]) /
                         dot3(pos_new[0], pos_new[1], pos_new[2], pos_new[0], pos_new[1], pos_new[2]));
            r[18] = sqrt(dot3(dv[0], dv[1], dv[2], dv[0], dv[1], dv[2]) /
                         dot3(vel_g[0], vel_g[1], vel_g[2], vel_g[0], vel_g[1], vel_g[2]));
            r[COLREC - 1] = (double)b;
        }
    }
    const double mass_rem = mass_imp + mass_tar;
    H.r_planet[I] = 0.;
    H.r_planet[T] *= cbrt_exact_ops(mass_rem / mass_tar);
    const int64_t tar_id = H.id[T];
    H.id[I] = tar_id;
    H.flags[I] |= GPLUM_FLAG_DEAD;
    for (int a = 0; a < n; a++)
        if (H.twin_of[b + a] == ic) { H.id[b + a] = tar_id; H.flags[b + a] |= GPLUM_FLAG_DEAD; }
    H.flags[T] |= GPLUM_FLAG_MERGED;
    for (int a = 0; a < n; a++) {
        if (a == ic || a == jc || H.twin_of[b + a] == ic || H.twin_of[b + a] == jc) {
            st3(H.pos, b + a, pos_new);
            st3(H.vel, b + a, vel_new);
        }
    }
    // calcEnergyDissipation (n_frag = 0)
    double e_kin = 0., s, vr[3];
    for (int k = 0; k < 3; k++) vr[k] = vel_imp[k] - vel_tar[k];
    s = 0.5 * mass_imp * mass_tar / (mass_imp + mass_tar);
    e_kin -= (s * vr[0]) * vr[0] + (s * vr[1]) * vr[1] + (s * vr[2]) * vr[2];
    for (int k = 0; k < 3; k++) vr[k] = vel_new[k] - vel_g[k];
    s = 0.5 * (mass_imp - 0.);
    e_kin += (s * vr[0]) * vr[0] + (s * vr[1]) * vr[1] + (s * vr[2]) * vr[2];
    s = 0.5 * mass_tar;
    e_kin += (s * vr[0]) * vr[0] + (s * vr[1]) * vr[1] + (s * vr[2]) * vr[2];
    double e_int = 0.;
    const double dr[3] = {pos_imp[0] - pos_tar[0], pos_imp[1] - pos_tar[1], pos_imp[2] - pos_tar[2]};
    const double dr2 = dot3(dr[0], dr[1], dr[2], dr[0], dr[1], dr[2]) + P.eps2;
    const double rinv = sqrt(1. / dr2);
    double dphi = H.mass[I] * rinv;
    e_int += H.mass[T] * dphi;
    for (int a = 0; a < n; a++) if (H.twin_of[b + a] == jc) e_int += H.mass[b + a] * dphi;
    for (int i2 = 0; i2 < n; i2++) {
        if (H.twin_of[b + i2] != ic) continue;
        dphi = H.mass[b + i2] * rinv;
        e_int += H.mass[T] * dphi;
        for (int a = 0; a < n; a++) if (H.twin_of[b + a] == jc) e_int += H.mass[b + a] * dphi;
    }
    double e_sun = 0.;
    e_sun += mass_imp / sqrt(dot3(pos_imp[0], pos_imp[1], pos_imp[2], pos_imp[0], pos_imp[1], pos_imp[2]) + P.eps2) +
             mass_tar / sqrt(dot3(pos_tar[0], pos_tar[1], pos_tar[2], pos_tar[0], pos_tar[1], pos_tar[2]) + P.eps2);
    double d2 = dot3(H.pos[3 * T], H.pos[3 * T + 1], H.pos[3 * T + 2], H.pos[3 * T], H.pos[3 * T + 1],
                     H.pos[3 * T + 2]) + P.eps2;
    e_sun -= mass_tar * sqrt(1. / d2);
    d2 = dot3(H.pos[3 * I], H.pos[3 * I + 1], H.pos[3 * I + 2], H.pos[3 * I], H.pos[3 * I + 1],
              H.pos[3 * I + 2]) + P.eps2;
    e_sun -= (mass_imp - 0.) * sqrt(1. / d2);
    e_sun *= P.m_sun;
    // merge_list update
    H.twin_of[I] = jc;
    if (H.flags[I] & GPLUM_FLAG_MERGED) {
        for (int a = 0; a < n; a++) if (H.twin_of[b + a] == ic) H.twin_of[b + a] = jc;
        H.flags[I] &= ~GPLUM_FLAG_MERGED;
    }
    return e_kin + e_int + e_sun;
}

// --------------------------------------------------------------- kernel

struct Reduce {              // block-level reductions through LDS
    double d[HERMITE_BIG_BLOCK];
    int i[HERMITE_BIG_BLOCK];
    int j[HERMITE_BIG_BLOCK];
};

__device__ inline double block_min(Reduce &R, double v) {
    const int t = threadIdx.x;
    __syncthreads();
    R.d[t] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w && R.d[t + w] < R.d[t]) R.d[t] = R.d[t + w];
        __syncthreads();
    }
    const double r = R.d[0];
    __syncthreads();
    return r;
}

// argmin of (R, i) lexicographic, j carried along; returns R, sets *bi, *bj
__device__ inline double block_argmin(Reduce &R, double v, int i, int j, int *bi, int *bj) {
    const int t = threadIdx.x;
    __syncthreads();
    R.d[t] = v; R.i[t] = i; R.j[t] = j;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w) {
            const double a = R.d[t], c = R.d[t + w];
            if (c < a || (c == a && R.i[t + w] < R.i[t])) {
                R.d[t] = c; R.i[t] = R.i[t + w]; R.j[t] = R.j[t + w];
            }
        }
        __syncthreads();
    }
    const double r = R.d[0];
    *bi = R.i[0];
    *bj = R.j[0];
    __syncthreads();
    return r;
}

// Minimum over the 64 lanes of the wave (every lane active; values not NaN): four DPP
// steps inside each row of 16 lanes, then the four row minima read as scalars.  Far
// shorter latency than a shuffle butterfly through the LDS crossbar.
__device__ inline double wave_min(double v) {
    v = dmin(v, dpp_f64<DPP_QUAD_1032>(v));
    v = dmin(v, dpp_f64<DPP_QUAD_2301>(v));
    v = dmin(v, dpp_f64<DPP_ROW_HALF_MIRROR>(v));
    v = dmin(v, dpp_f64<DPP_ROW_MIRROR>(v));
    return dmin(dmin(readlane_f64(v, 0), readlane_f64(v, 16)), dmin(readlane_f64(v, 32), readlane_f64(v, 48)));
}
// two minima at once (independent chains)
__device__ inline void wave_min2(double &u, double &v) {
    u = dmin(u, dpp_f64<DPP_QUAD_1032>(u)); v = dmin(v, dpp_f64<DPP_QUAD_1032>(v));
    u = dmin(u, dpp_f64<DPP_QUAD_2301>(u)); v = dmin(v, dpp_f64<DPP_QUAD_2301>(v));
    u = dmin(u, dpp_f64<DPP_ROW_HALF_MIRROR>(u)); v = dmin(v, dpp_f64<DPP_ROW_HALF_MIRROR>(v));
    u = dmin(u, dpp_f64<DPP_ROW_MIRROR>(u)); v = dmin(v, dpp_f64<DPP_ROW_MIRROR>(v));
    u = dmin(dmin(readlane_f64(u, 0), readlane_f64(u, 16)), dmin(readlane_f64(u, 32), readlane_f64(u, 48)));
    v = dmin(dmin(readlane_f64(v, 0), readlane_f64(v, 16)), dmin(readlane_f64(v, 32), readlane_f64(v, 48)));
}
__device__ inline void wave_argmin(double &v, int &i, int &j) {
    for (int d = 32; d > 0; d >>= 1) {
        const double ov = __shfl_xor(v, d);
        const int oi = __shfl_xor(i, d), oj = __shfl_xor(j, d);
        if (ov < v || (ov == v && oi < i)) { v = ov; i = oi; j = oj; }
    }
}

// Orders the wave's own global-memory writes before its later reads (lane 0 writes
// state that all lanes read back): the wave's memory operations are complete at the
// fence, and the barrier keeps the compiler from moving accesses across it.
__device__ inline void wave_sync() {
    __threadfence_block();
    __builtin_amdgcn_wave_barrier();
}

// The active-particle correction of integrate_from for clusters above HARD_SEQ_MAX
// particles (H.act marks the active ones): the k-th active particle (ascending index)
// goes to wave k mod (waves per workgroup), and the wave's 64 lanes evaluate its hard
// force together (coop_pair_sum); every lane carries the same particle state in
// registers and lane 0 writes it back.  Active particles only read the predicted
// j-data, so the waves need not wait for each other.  Twins of a merged target are
// integrated alongside with their own cut-off radius, as in the per-lane path.  tw: the
// cluster's twin_of (cluster-local indices; an LDS copy where the kernel keeps one).
__device__ void coop_correct_active(const HP &P, HardView &H, const JData &J, int b, int n, double t_s,
                                    int flag_col, long long &pairs, int wave, int n_wave, const int32_t *tw) {
    const int lane = (int)(threadIdx.x & 63);
    const bool lead = lane == 0;
    int rank = 0;                          // active particles seen so far (wave-uniform)
    for (int i0 = 0; i0 < n; i0 += 64) {
    unsigned long long mask = __ballot(i0 + lane < n && H.act[b + i0 + lane] != 0);
    while (mask) {
        const int i = i0 + __ffsll(mask) - 1;
        mask &= mask - 1;
        if ((rank++) % n_wave != wave) continue;
        const int g = b + i;
        const bool merged = (H.flags[g] & GPLUM_FLAG_MERGED) != 0;
        if (lead) pairs += 3 * (long long)(n - 1);
        // state of g (predicted by the predict pass)
        double xp[3], vp[3], a0s[3], j0s[3], a0d[3], j0d[3], as[3], js[3], ad[3], jd[3];
        double x[3], v[3], a2s[3], a3s[3], a2d[3], a3d[3];
        ld3(H.xp, g, xp); ld3(H.vp, g, vp);
        ld3(H.a0_s, g, a0s); ld3(H.j0_s, g, j0s); ld3(H.a0_d, g, a0d); ld3(H.j0_d, g, j0d);
        const double rinv = H.rinv[g];
        const int64_t id = H.id[g];
        star_acc_jerk(P, xp, vp, as, js, nullptr);
        coop_pair_sum(P, J, n, id, xp, vp, rinv, ad, jd);
        if (merged) {
            // mergeAccJerk over the target and its twins (twin index order)
            double sa[3], sj[3], mass = H.mass[g];
            for (int k = 0; k < 3; k++) { sa[k] = H.mass[g] * ad[k]; sj[k] = H.mass[g] * jd[k]; }
            for (int j = 0; j < n; j++) {
                if (tw[j] != i) continue;
                double txp[3], tvp[3], tad[3], tjd[3], tas[3], tjs[3];
                ld3(H.xp, b + j, txp); ld3(H.vp, b + j, tvp);
                star_acc_jerk(P, txp, tvp, tas, tjs, nullptr);
                coop_pair_sum(P, J, n, H.id[b + j], txp, tvp, H.rinv[b + j], tad, tjd);
                if (lead) { st3(H.acc_s, b + j, tas); st3(H.jerk_s, b + j, tjs); pairs += 3 * (long long)(n - 1); }
                for (int k = 0; k < 3; k++) { sa[k] += H.mass[b + j] * tad[k]; sj[k] += H.mass[b + j] * tjd[k]; }
                mass += H.mass[b + j];
            }
            const double minv = 1. / mass;
            for (int k = 0; k < 3; k++) { ad[k] = sa[k] * minv; jd[k] = sj[k] * minv; }
        }
        const double t_c = t_s - H.time[g];
        for (int ite = 0; ite < 2; ite++) {
            correct_regs(t_c, xp, vp, as, a0s, j0s, js, ad, a0d, j0d, jd, x, v, a2s, a3s, a2d, a3d);
            if (merged && lead)
                for (int j = 0; j < n; j++)
                    if (tw[j] == i) {
                        // the twin's correction from its own predicted state and the merged acc/jerk
                        double txp[3], tvp[3], ta0s[3], tj0s[3], ta0d[3], tj0d[3], tas[3], tjs[3];
                        double tx[3], tv[3], t2s[3], t3s[3], t2d[3], t3d[3];
                        ld3(H.xp, b + j, txp); ld3(H.vp, b + j, tvp);
                        ld3(H.a0_s, b + j, ta0s); ld3(H.j0_s, b + j, tj0s);
                        ld3(H.a0_d, b + j, ta0d); ld3(H.j0_d, b + j, tj0d);
                        ld3(H.acc_s, b + j, tas); ld3(H.jerk_s, b + j, tjs);
                        correct_regs(t_c, txp, tvp, tas, ta0s, tj0s, tjs, ad, ta0d, tj0d, jd, tx, tv, t2s, t3s, t2d,
                                     t3d);
                        st3(H.pos, b + j, tx); st3(H.vel, b + j, tv);
                    }
            wave_sync();
            star_acc_jerk(P, x, v, as, js, nullptr);
            coop_pair_sum(P, J, n, id, x, v, rinv, ad, jd);
            if (merged) {
                double sa[3], sj[3], mass = H.mass[g];
                for (int k = 0; k < 3; k++) { sa[k] = H.mass[g] * ad[k]; sj[k] = H.mass[g] * jd[k]; }
                for (int j = 0; j < n; j++) {
                    if (tw[j] != i) continue;
                    double tx[3], tv[3], tad[3], tjd[3], tas[3], tjs[3];
                    ld3(H.pos, b + j, tx); ld3(H.vel, b + j, tv);
                    star_acc_jerk(P, tx, tv, tas, tjs, nullptr);
                    coop_pair_sum(P, J, n, H.id[b + j], tx, tv, H.rinv[b + j], tad, tjd);
                    wave_sync();
                    if (lead) { st3(H.acc_s, b + j, tas); st3(H.jerk_s, b + j, tjs); }
                    for (int k = 0; k < 3; k++) { sa[k] += H.mass[b + j] * tad[k]; sj[k] += H.mass[b + j] * tjd[k]; }
                    mass += H.mass[b + j];
                }
                const double minv = 1. / mass;
                for (int k = 0; k < 3; k++) { ad[k] = sa[k] * minv; jd[k] = sj[k] * minv; }
            }
        }
        double time = H.time[g] + t_c;
        double dt = H.dt[g];
        if (!flag_col) dt = dt_block(P, time, dt, H.acc0[g], ad, jd, a2d, a3d, as, js, a2s, a3s);
        wave_sync();
        if (lead) {
            st3(H.pos, g, x); st3(H.vel, g, v);
            st3(H.acc_s, g, as); st3(H.jerk_s, g, js); st3(H.acc_d, g, ad); st3(H.jerk_d, g, jd);
            st3(H.a2_s, g, a2s); st3(H.a3_s, g, a3s); st3(H.a2_d, g, a2d); st3(H.a3_d, g, a3d);
            H.time[g] = time;
            if (!flag_col) H.dt[g] = dt;
            if (merged)
                for (int j = 0; j < n; j++)
                    if (tw[j] == i) {
                        st3(H.acc_d, b + j, ad); st3(H.jerk_d, b + j, jd);
                        H.time[b + j] += t_c;
                        if (!flag_col) H.dt[b + j] = dt;
                    }
        }
        wave_sync();
    }
    }
}

// ---------------------------- multi-wave kernels: lane groups and wave-level reductions

// Per-workgroup scratch of the multi-wave kernels' block steps: one slot per wave for
// each reduction (separate slots, so that a reduction needs no barrier before its
// writes), and the list of the block step's active particles.
struct WaveSlots {
    double ts[8], sys[8], rm[8];
    int lo[8], hi[8];
    int na, nm;                // active particles without twins (listed in act), merged targets
    int act[HERMITE_LDS];
    int32_t tw[HERMITE_LDS];   // the cluster's twin_of, copied at each block step's prediction
    // merged targets of the block step, their twins (concatenated, ascending index per
    // target) and the lane-group layout of the bodies to integrate: slot k holds body
    // ent[k] (-1: empty), the first slot of its family efam[k] and the family's size
    // efn[k] (FAM_MERGED set for a merged family; 1 for a body without twins)
    int mt[HERMITE_LDS];
    int twl[HERMITE_LDS];
    int twn[HERMITE_LDS];
    int ent[2 * HERMITE_LDS], efam[2 * HERMITE_LDS], efn[2 * HERMITE_LDS];
    int nslot, gs, nfam;       // slots in use, log2 of the lanes per slot, families placed
};
constexpr int FAM_MERGED = 1 << 16;
constexpr int FAM_MAX = 8;     // largest family integrated in the lane groups (64 / 8 lanes)
constexpr int FAM_GROUPS = 16; // families per block step in the lane groups (slot capacity)

// Minimum over the workgroup (every lane active, values not NaN): the wave's minimum by
// DPP, the waves' minima through slot[] and one barrier.  Exact, hence the value of
// block_min.
__device__ inline double block_min_waves(double v, double *slot) {
    v = wave_min(v);
    if ((threadIdx.x & 63) == 0) slot[threadIdx.x >> 6] = v;
    __syncthreads();
    double m = slot[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = dmin(m, slot[w]);
    return m;
}

// (R, lo, hi) lexicographic minimum over the workgroup (every lane active).
__device__ inline bool rlh_less(double r, int lo, int hi, double r2, int lo2, int hi2) {
    return r < r2 || (r == r2 && (lo < lo2 || (lo == lo2 && hi < hi2)));
}
__device__ inline void block_argmin_waves(double &v, int &lo, int &hi, WaveSlots &W) {
    for (int d = 32; d > 0; d >>= 1) {
        const double ov = __shfl_xor(v, d);
        const int ol = __shfl_xor(lo, d), oh = __shfl_xor(hi, d);
        if (rlh_less(ov, ol, oh, v, lo, hi)) { v = ov; lo = ol; hi = oh; }
    }
    if ((threadIdx.x & 63) == 0) {
        const int w = threadIdx.x >> 6;
        W.rm[w] = v; W.lo[w] = lo; W.hi[w] = hi;
    }
    __syncthreads();
    v = W.rm[0]; lo = W.lo[0]; hi = W.hi[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++)
        if (rlh_less(W.rm[w], W.lo[w], W.hi[w], v, lo, hi)) { v = W.rm[w]; lo = W.lo[w]; hi = W.hi[w]; }
}

// Subtree sums of coop_pair_sum's butterfly.  With all 64 lanes on one particle, lane l
// holds the partial P_l (pairs j = l, l + 64, ... in ascending j) and the xor butterfly
// 32, 16, ..., 1 adds them: after the level of distance 2^k every lane holds V_k(l mod
// 2^k), the sum of the partials m = l (mod 2^k), where V_k(r) = V_(k+1)(r) + V_(k+1)(r +
// 2^k).  A group of S = 2^s lanes holding one particle computes the same sums: lane r of
// the group builds V_s(r) from its 64/S partials m = r + S bitrev(q), q = 0, 1, ... (the
// leaves of that subtree from left to right, so that a binary counter over q adds two
// sibling subtrees as soon as the second one is complete), then the group runs the
// butterfly's last s levels.  The same additions of the same values as coop_pair_sum,
// hence the same force bit for bit.  K = ceil(n / 64) pairs per partial, evaluated side
// by side (partials with fewer pairs evaluate a clamped j and drop it).  Every lane of
// the wave calls it (groups of invalid particles compute a stand-in); 3 <= s <= 6.
template <int K>
__device__ inline void group_pair_sum(const HP &P, const JData &J, int n, int s, int r, int64_t idi,
                                      const double xi[3], const double vi[3], double rinv_i, double acc[3],
                                      double jerk[3]) {
    const int depth = 6 - s;                      // levels of the lane's own subtree
    const int nq = 1 << depth;
    double st[3][6];
    double c[6];
    for (int q = 0; q < nq; q++) {
        const int m = r + ((depth ? (int)(__builtin_bitreverse32((unsigned)q) >> (32 - depth)) : 0) << s);
        for (int d = 0; d < 6; d++) c[d] = 0.;
#pragma unroll
        for (int u0 = 0; u0 < K; u0 += 2) {
            constexpr int L = 2;
            double ta[L][3], tj[L][3];
            int jj[L];
#pragma unroll
            for (int u = 0; u < L; u++) {
                jj[u] = min(m + 64 * (u0 + u), n - 1);
                if (u0 + u < K) pair_term_j(P, J, jj[u], xi, vi, rinv_i, ta[u], tj[u], nullptr);
            }
#pragma unroll
            for (int u = 0; u < L; u++) {
                if (u0 + u >= K) continue;
                const bool use = m + 64 * (u0 + u) < n && J.id[jj[u]] != idi;
                for (int d = 0; d < 3; d++) {
                    c[d] = use ? c[d] + ta[u][d] : c[d];
                    c[3 + d] = use ? c[3 + d] + tj[u][d] : c[3 + d];
                }
            }
        }
        // binary counter: level k holds the left subtree of 2^k leaves until its sibling
        // is complete
        bool stored = false;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (stored || k >= depth) continue;
            if (q & (1 << k)) {
                for (int d = 0; d < 6; d++) c[d] = st[k][d] + c[d];
            } else {
                for (int d = 0; d < 6; d++) st[k][d] = c[d];
                stored = true;
            }
        }
    }
    // the last s levels of the butterfly, inside the group (as in coop_pair_sum)
    if (s == 6)
        for (int d = 0; d < 6; d++) c[d] += __shfl_xor(c[d], 32);
    if (s >= 5)
        for (int d = 0; d < 6; d++) c[d] += __shfl_xor(c[d], 16);
    if (s >= 4)
        for (int d = 0; d < 6; d++) c[d] += dpp_f64<DPP_ROW_ROR8>(c[d]);
    for (int d = 0; d < 6; d++) c[d] += dpp_f64<DPP_QUAD_3210>(dpp_f64<DPP_ROW_HALF_MIRROR>(c[d]));
    for (int d = 0; d < 6; d++) c[d] += dpp_f64<DPP_QUAD_2301>(c[d]);
    for (int d = 0; d < 6; d++) c[d] += dpp_f64<DPP_QUAD_1032>(c[d]);
    for (int d = 0; d < 3; d++) { acc[d] = c[d]; jerk[d] = c[3 + d]; }
}

// The active-particle correction of integrate_from (hermite.h:300-375) in lane groups,
// for clusters with j-data in LDS.  The bodies to integrate are laid out in slots
// (WaveSlots::ent): the active particles without twins, and merged targets with their
// twins as families of consecutive slots inside one wave (a merged target integrates its
// twins alongside, each with its own position and cut-off radius, and mergeAccJerk
// averages their hard forces by mass).  A slot gets S = 2^s lanes (W.gs); each group
// evaluates its body's hard force by group_pair_sum, and the members of a family add
// their forces by shuffles in the order of coop_correct_active (target first, then the
// twins in ascending index).  The steps and their order are those of
// coop_correct_active, so the results are too.  Every thread of the workgroup calls it
// (no barrier inside).
template <int K>
__device__ void group_correct_active(const HP &P, HardView &H, const JData &J, int b, int n, double t_s,
                                     int flag_col, long long &pairs, const WaveSlots &W,
                                     const double *__restrict__ a0_lds) {
    const int t = threadIdx.x;
    const int B = blockDim.x;
    const int s = W.gs, E = 64 >> s;
    const int r = t & ((1 << s) - 1);
    const int nslot = W.nslot;
    const bool fams = W.nfam > 0;
    const int per_round = B >> s;
    const int wave_first = (t & ~63) >> s;     // the wave's first slot in a round
    for (int a0 = 0; a0 < nslot; a0 += per_round) {
        if (a0 + wave_first >= nslot) continue;   // no body for this wave (wave-uniform)
        const int slot = a0 + (t >> s);
        const int e = slot < nslot ? W.ent[slot] : -1;
        const bool valid = e >= 0;
        const int g = b + (valid ? e : 0);      // empty slots compute a stand-in
        int fn = 1, ff = slot;
        bool mfam = false;
        if (fams && slot < nslot) {
            fn = W.efn[slot] & (FAM_MERGED - 1);
            mfam = (W.efn[slot] & FAM_MERGED) != 0;
            ff = W.efam[slot];
        }
        const bool twin = ff != slot;
        const int gt = twin ? b + W.ent[ff] : g;  // the family's target
        const int ffw = ff - (a0 + wave_first);   // its slot within the wave
        const bool wave_fam = fams && __ballot(mfam) != 0ull;
        const bool lead = valid && r == 0;
        const double mass_e = H.mass[g];
        double as[3], js[3], ad[3], jd[3], x[3], v[3];
        // mergeAccJerk over the family (hermite.h:36-74), from the members' lanes
        auto merge = [&]() __attribute__((always_inline)) {
            if (!wave_fam) return;
            double sa[6] = {0., 0., 0., 0., 0., 0.}, msum = 0.;
            for (int k = 0; k < E; k++) {
                const int src = ((ffw + k) << s) + r;
                const double mk = __shfl(mass_e, src);
                double ak[6];
                for (int d = 0; d < 3; d++) { ak[d] = __shfl(ad[d], src); ak[3 + d] = __shfl(jd[d], src); }
                if (k < fn) {
                    if (k == 0) {
                        for (int d = 0; d < 6; d++) sa[d] = mk * ak[d];
                        msum = mk;
                    } else {
                        for (int d = 0; d < 6; d++) sa[d] += mk * ak[d];
                        msum += mk;
                    }
                }
            }
            if (mfam) {
                const double minv = 1. / msum;
                for (int d = 0; d < 3; d++) { ad[d] = sa[d] * minv; jd[d] = sa[3 + d] * minv; }
            }
        };
        // Registers are scarce here (the pair sums keep a subtree stack per lane): the
        // body's predicted state and its acc/jerk at the start of the block step are
        // reread from LDS (the j-data and a0) for each correction instead of being held
        // across the pair sums, and the correction's a2/a3 terms go through H.
        const int el = valid ? e : 0;
        const double rinv = H.rinv[g];
        const int64_t id = H.id[g];
        {
            double xp[3], vp[3];
            ld3(J.x, el, xp); ld3(J.v, el, vp);
            group_pair_sum<K>(P, J, n, s, r, id, xp, vp, rinv, ad, jd);
            star_acc_jerk(P, xp, vp, as, js, nullptr);
        }
        merge();
        const double t_c = t_s - H.time[gt];    // twins carry their target's time
        for (int ite = 0; ite < 2; ite++) {
            {
                double xp[3], vp[3], a0s[3], j0s[3], a0d[3], j0d[3], a2s[3], a3s[3], a2d[3], a3d[3];
                ld3(J.x, el, xp); ld3(J.v, el, vp);
                const double *q = a0_lds + 12 * el;
                for (int k = 0; k < 3; k++) {
                    a0s[k] = q[k]; j0s[k] = q[3 + k]; a0d[k] = q[6 + k]; j0d[k] = q[9 + k];
                }
                correct_regs(t_c, xp, vp, as, a0s, j0s, js, ad, a0d, j0d, jd, x, v, a2s, a3s, a2d, a3d);
                if (ite == 1 && lead && !twin) {
                    st3(H.a2_s, g, a2s); st3(H.a3_s, g, a3s); st3(H.a2_d, g, a2d); st3(H.a3_d, g, a3d);
                }
            }
            star_acc_jerk(P, x, v, as, js, nullptr);
            group_pair_sum<K>(P, J, n, s, r, id, x, v, rinv, ad, jd);
            merge();
        }
        wave_sync();     // the leader's a2/a3 before the group reads them back
        const double time = H.time[g] + t_c;
        double dt = H.dt[g];
        if (!flag_col) {
            double a2s[3], a3s[3], a2d[3], a3d[3];
            ld3(H.a2_s, g, a2s); ld3(H.a3_s, g, a3s); ld3(H.a2_d, g, a2d); ld3(H.a3_d, g, a3d);
            dt = dt_block(P, time, dt, H.acc0[g], ad, jd, a2d, a3d, as, js, a2s, a3s);
        }
        if (wave_fam) {                           // twins take their target's step
            const double dtt = __shfl(dt, (ffw << s) + r);
            if (twin) dt = dtt;
        }
        if (lead) {
            pairs += 3 * (long long)(n - 1);
            st3(H.pos, g, x); st3(H.vel, g, v);
            st3(H.acc_s, g, as); st3(H.jerk_s, g, js); st3(H.acc_d, g, ad); st3(H.jerk_d, g, jd);
            H.time[g] = time;
            if (!flag_col) H.dt[g] = dt;
        }
    }
}

// After the collision test found the pair (ic, jc) in a block step and the active
// particles were corrected (hermite.h:289-296, 449-458): perfect merging, the full hard
// force of everybody, the merged targets' acc/jerk, the initial block step of the live
// particles and the twins' steps.  Every thread of one workgroup calls it.
__device__ void collision_tail(const HP &P, HardView &H, int b, int n, int ic, int jc, double &ed,
                               long long &pairs) {
    const int t = threadIdx.x;
    const int B = blockDim.x;
    if (t == 0) ed += perfect_merge(P, H, b, n, ic, jc);
    __threadfence_block();
    __syncthreads();
    for (int i = t; i < n; i += B) { gravity_full(P, H, b, n, i); pairs += n - 1; }
    __threadfence_block();
    __syncthreads();
    for (int i = t; i < n; i += B)
        if (H.flags[b + i] & GPLUM_FLAG_MERGED) merge_acc_jerk(H, b, n, i);
    __threadfence_block();
    __syncthreads();
    for (int i = t; i < n; i += B) {
        const int g = b + i;
        if (!(H.flags[g] & GPLUM_FLAG_DEAD)) {
            double ad[3], jd[3], as[3], js[3];
            ld3(H.acc_d, g, ad); ld3(H.jerk_d, g, jd); ld3(H.acc_s, g, as); ld3(H.jerk_s, g, js);
            H.dt[g] = dt_initial(P, H.time[g], H.dt[g], H.acc0[g], ad, jd, as, js);
        }
    }
    __threadfence_block();
    __syncthreads();
    for (int i = t; i < n; i += B)
        if (H.twin_of[b + i] >= 0) H.dt[b + i] = H.dt[b + H.twin_of[b + i]];
    __threadfence_block();
    __syncthreads();
}

// Block steps of one cluster from (sys_time, loop) to the end of the soft step on the
// global-memory state H (j-data through J), then the final calcGravity.  Every
// thread of the workgroup calls it; collisions and twins handled here.  COOP: the
// cluster has more than HARD_SEQ_MAX particles (the other correction path is left out
// of the instantiation, which keeps its register use down).
template <bool LDSJ, bool COOP>
__device__ void integrate_from(const HP &P, HardView &H, const JData &J, Reduce &R, double *s_x, double *s_v,
                               int64_t *s_id, double *s_fr, int b, int n, double &sys_time, long long &loop,
                               double &ed, int &n_col, long long &pairs, WaveSlots *W = nullptr,
                               double *s_a0 = nullptr) {
    const int t = threadIdx.x;
    const int B = blockDim.x;
    const double time_end = P.dt_tree;
    while (sys_time < time_end && loop < BLOCK_STEP_CAP) {
        // makeActiveList: t_s = min(time + dt) over live particles
        double tl = 1.e10;
        for (int i = t; i < n; i += B)
            if (!(H.flags[b + i] & GPLUM_FLAG_DEAD)) tl = dmin(tl, H.time[b + i] + H.dt[b + i]);
        const double t_s = COOP ? block_min_waves(tl, W->ts) : block_min(R, tl);
        // predict all (hermite.h:277-283)
        for (int i = t; i < n; i += B) {
            predict(H, b + i, t_s - H.time[b + i]);
            if (LDSJ) {
                for (int k = 0; k < 3; k++) {
                    s_x[3 * i + k] = H.xp[3 * (b + i) + k];
                    s_v[3 * i + k] = H.vp[3 * (b + i) + k];
                }
                s_id[i] = H.id[b + i];
                // f R of the live particles for the collision test (-1: dead)
                s_fr[i] = (H.flags[b + i] & GPLUM_FLAG_DEAD) ? -1. : H.f[b + i] * H.r_planet[b + i];
                if (COOP) {
                    W->tw[i] = H.twin_of[b + i];
                    // acc/jerk at the start of the block step (predict's a0), for the lane groups
                    for (int k = 0; k < 3; k++) {
                        s_a0[12 * i + k] = H.a0_s[3 * (b + i) + k];
                        s_a0[12 * i + 3 + k] = H.j0_s[3 * (b + i) + k];
                        s_a0[12 * i + 6 + k] = H.a0_d[3 * (b + i) + k];
                        s_a0[12 * i + 9 + k] = H.j0_d[3 * (b + i) + k];
                    }
                }
            }
        }
        __threadfence_block();
        __syncthreads();
        // collisionDetermination (hermite.h:98-196)
        int flag_col = 0, ic = -1, jc = -1;
        if (P.collision) {
            double rmin = 1.;
            int bi = 0x7fffffff, bj = COOP ? 0x7fffffff : -1;
            if (LDSJ && COOP) {
                // the unordered pairs {a, c} over all threads: c = (a + d) mod n for d = 1 ..
                // (n - 1) / 2 and every a, and for even n d = n / 2 with a < n / 2.  R is
                // symmetric bit for bit (squared differences, a commutative sum of f R),
                // so the (R, lower, higher index) minimum is the (R, i, first j) minimum
                // of the ordered loop below.
                // Thread t takes particle a = t mod n and the r-th of R = B / n runs of
                // the distances d (a difference of signs leaves d2 unchanged).
                const int R = max(1, B / n);
                if (t < n * R) {
                    const int a = t % n, r = t / n;
                    const int h = (n - 1) >> 1;
                    const int dlen = (h + R - 1) / R;
                    const int d0 = 1 + r * dlen, d1 = min(h, d0 + dlen - 1);
                    const double fra = s_fr[a];
                    const double xa = s_x[3 * a], ya = s_x[3 * a + 1], za = s_x[3 * a + 2];
                    auto pair = [&](int c) {
                        const double frc = s_fr[c];
                        if (frc < 0.) return;
                        const double dx = xa - s_x[3 * c];
                        const double dy = ya - s_x[3 * c + 1];
                        const double dz = za - s_x[3 * c + 2];
                        const double d2 = dot3(dx, dy, dz, dx, dy, dz);
                        const double r2 = fra + frc;
                        if (!(d2 < COL_FILTER * r2 * r2)) return;
                        const double Rr = sqrt(d2) / r2;
                        const int lo = min(a, c), hi = max(a, c);
                        if (Rr < 1. && rlh_less(Rr, lo, hi, rmin, bi, bj)) { rmin = Rr; bi = lo; bj = hi; }
                    };
                    if (fra >= 0.) {
                        // four distances at a time (their LDS reads overlap); the minimum
                        // does not depend on the order in which pairs are met
                        int d = d0;
                        for (; d + 3 <= d1; d += 4) {
                            double frc[4], ddx[4], ddy[4], ddz[4];
                            int cc[4];
#pragma unroll
                            for (int u = 0; u < 4; u++) {
                                cc[u] = a + d + u < n ? a + d + u : a + d + u - n;
                                frc[u] = s_fr[cc[u]];
                                ddx[u] = xa - s_x[3 * cc[u]];
                                ddy[u] = ya - s_x[3 * cc[u] + 1];
                                ddz[u] = za - s_x[3 * cc[u] + 2];
                            }
#pragma unroll
                            for (int u = 0; u < 4; u++) {
                                if (frc[u] < 0.) continue;
                                const double d2 = dot3(ddx[u], ddy[u], ddz[u], ddx[u], ddy[u], ddz[u]);
                                const double r2 = fra + frc[u];
                                if (!(d2 < COL_FILTER * r2 * r2)) continue;
                                const double Rr = sqrt(d2) / r2;
                                const int lo = min(a, cc[u]), hi = max(a, cc[u]);
                                if (Rr < 1. && rlh_less(Rr, lo, hi, rmin, bi, bj)) { rmin = Rr; bi = lo; bj = hi; }
                            }
                        }
                        for (; d <= d1; d++) pair(a + d < n ? a + d : a + d - n);
                        // even n: the pairs at distance n / 2, once each (a < n / 2)
                        if (!(n & 1) && r == R - 1 && a < (n >> 1)) pair(a + (n >> 1));
                    }
                }
            } else
            for (int i = t; i < n; i += B) {
                if (LDSJ) {
                    // predicted positions and f R from LDS
                    const double fri = s_fr[i];
                    if (fri < 0.) continue;
                    const double xi = s_x[3 * i], yi = s_x[3 * i + 1], zi = s_x[3 * i + 2];
                    for (int j = 0; j < n; j++) {
                        const double frj = s_fr[j];
                        if (i == j || frj < 0.) continue;
                        const double dx = xi - s_x[3 * j];
                        const double dy = yi - s_x[3 * j + 1];
                        const double dz = zi - s_x[3 * j + 2];
                        const double d2 = dot3(dx, dy, dz, dx, dy, dz);
                        const double r2 = fri + frj;
                        if (!(d2 < COL_FILTER * r2 * r2)) continue;
                        const double Rr = sqrt(d2) / r2;
                        if (Rr < rmin) { rmin = Rr; bi = i; bj = j; }
                    }
                } else {
                    if (H.flags[b + i] & GPLUM_FLAG_DEAD) continue;
                    for (int j = 0; j < n; j++) {
                        if (i == j || (H.flags[b + j] & GPLUM_FLAG_DEAD)) continue;
                        const double dx = H.xp[3 * (b + i)] - H.xp[3 * (b + j)];
                        const double dy = H.xp[3 * (b + i) + 1] - H.xp[3 * (b + j) + 1];
                        const double dz = H.xp[3 * (b + i) + 2] - H.xp[3 * (b + j) + 2];
                        const double d2 = dot3(dx, dy, dz, dx, dy, dz);
                        const double r2 = H.f[b + i] * H.r_planet[b + i] + H.f[b + j] * H.r_planet[b + j];
                        if (!(d2 < COL_FILTER * r2 * r2)) continue;
                        const double Rr = sqrt(d2) / r2;
                        if (Rr < rmin) { rmin = Rr; bi = i; bj = j; }
                    }
                }
            }
            int gi = bi, gj = bj;
            double rm = rmin;
            if (COOP) block_argmin_waves(rm, gi, gj, *W);
            else rm = block_argmin(R, rmin, bi, bj, &gi, &gj);
            if (rm < 1.) {
                flag_col = 1;
                if (H.mass[b + gi] < H.mass[b + gj]) { ic = gi; jc = gj; } else { ic = gj; jc = gi; }
            }
            if (flag_col)
                for (int i = t; i < n; i += B) H.dt[b + i] = t_s - H.time[b + i];
            __threadfence_block();
            __syncthreads();
        }
        // correct the active particles (hermite.h:300-375)
        if (COOP && LDSJ) {
            // The bodies of the block step, laid out by the first wave for the lane groups
            // (group_correct_active): the merged targets with their twins as families of
            // consecutive slots that stay inside one wave (at most FAM_MAX members and
            // FAM_GROUPS families; larger or further ones are flagged in H.act for
            // coop_correct_active), then the active particles without twins in ascending
            // index.  The two paths write disjoint particles and read only the predicted
            // j-data.
            if (t < 64) {
                const unsigned long long below = (1ull << t) - 1ull;
                int base = 0, nm = 0;
                for (int i0 = 0; i0 < n; i0 += 64) {
                    const int i = i0 + t;
                    int a = 0, m = 0;
                    if (i < n) {
                        const int f = H.flags[b + i];
                        a = !(f & GPLUM_FLAG_DEAD) && (flag_col || H.time[b + i] + H.dt[b + i] == t_s);
                        m = a && (f & GPLUM_FLAG_MERGED);
                        H.act[b + i] = 0;
                        a = a && !m;
                    }
                    const unsigned long long ma = __ballot(a), mm = __ballot(m);
                    if (a) W->act[base + __popcll(ma & below)] = i;
                    if (m) W->mt[nm + __popcll(mm & below)] = i;
                    base += __popcll(ma);
                    nm += __popcll(mm);
                }
                // each merged target's twins in ascending index
                int ntw = 0, nfam = 0, nbig = 0, fmem = 0, fmax = 1;
                for (int k = 0; k < nm; k++) {
                    const int tg = W->mt[k];
                    int c = 0;
                    for (int i0 = 0; i0 < n; i0 += 64) {
                        const int i = i0 + t;
                        const bool q = i < n && W->tw[i] == tg;
                        const unsigned long long qm = __ballot(q);
                        if (q) W->twl[ntw + c + __popcll(qm & below)] = i;
                        c += __popcll(qm);
                    }
                    if (1 + c > FAM_MAX || nfam >= FAM_GROUPS) {
                        if (t == 0) H.act[b + tg] = 1;     // coop_correct_active
                        c = -1;
                        nbig++;
                    } else {
                        ntw += c;
                        nfam++;
                        fmem += 1 + c;
                        fmax = max(fmax, 1 + c);
                    }
                    if (t == 0) W->twn[k] = c;
                }
                // lanes per slot: the most that hold every body at once (at least 8, and
                // few enough that a wave holds the largest family)
                const int ne = base + fmem;
                int sg = 6;
                while (sg > 3 && (ne * (1 << sg) > B || (64 >> sg) < fmax)) sg--;
                const int E = 64 >> sg;
                int pos = 0;
                if (t == 0) {
                    int tw0 = 0;
                    for (int k = 0; k < nm; k++) {
                        const int c = W->twn[k];
                        if (c < 0) continue;
                        const int fsz = 1 + c;
                        if ((pos % E) + fsz > E)
                            while (pos % E) { W->ent[pos] = -1; W->efam[pos] = pos; W->efn[pos] = 1; pos++; }
                        W->ent[pos] = W->mt[k];
                        for (int q = 0; q < c; q++) W->ent[pos + 1 + q] = W->twl[tw0 + q];
                        for (int q = 0; q < fsz; q++) { W->efam[pos + q] = pos; W->efn[pos + q] = fsz | FAM_MERGED; }
                        tw0 += c;
                        pos += fsz;
                    }
                }
                pos = __shfl(pos, 0);
                for (int k = t; k < base; k += 64) {
                    W->ent[pos + k] = W->act[k];
                    W->efam[pos + k] = pos + k;
                    W->efn[pos + k] = 1;
                }
                if (t == 0) { W->nslot = pos + base; W->gs = sg; W->nfam = nfam; W->na = base; W->nm = nbig; }
            }
            __threadfence_block();
            __syncthreads();
            const int K = (n + 63) >> 6;
            if (K <= 2) group_correct_active<2>(P, H, J, b, n, t_s, flag_col, pairs, *W, s_a0);
            else if (K == 3) group_correct_active<3>(P, H, J, b, n, t_s, flag_col, pairs, *W, s_a0);
            else group_correct_active<4>(P, H, J, b, n, t_s, flag_col, pairs, *W, s_a0);
            if (W->nm > 0)
                coop_correct_active(P, H, J, b, n, t_s, flag_col, pairs, (int)(threadIdx.x >> 6), B >> 6, W->tw);
        } else if (COOP) {
            for (int i = t; i < n; i += B)
                H.act[b + i] = !(H.flags[b + i] & GPLUM_FLAG_DEAD) &&
                               (flag_col || H.time[b + i] + H.dt[b + i] == t_s);
            __threadfence_block();
            __syncthreads();
            coop_correct_active(P, H, J, b, n, t_s, flag_col, pairs, (int)(threadIdx.x >> 6), B >> 6,
                                H.twin_of + b);
        } else
        for (int i = t; i < n; i += B) {
            const int g = b + i;
            if (H.flags[g] & GPLUM_FLAG_DEAD) continue;
            if (!flag_col && H.time[g] + H.dt[g] != t_s) continue;
            // a merged target integrates its dead twins alongside (hermite.h:300-375): each
            // twin feels the hard force with its own r_out (cut-off), and mergeAccJerk
            // averages target and twins by mass
            const bool merged = (H.flags[g] & GPLUM_FLAG_MERGED) != 0;
            pairs += 3 * (long long)(n - 1);
            gravity_pc(P, H, J, n, g, false);
            if (merged) {
                for (int j = 0; j < n; j++)
                    if (H.twin_of[b + j] == i) { gravity_pc(P, H, J, n, b + j, false); pairs += 3 * (long long)(n - 1); }
                merge_acc_jerk(H, b, n, i);
            }
            const double t_c = t_s - H.time[g];
            for (int ite = 0; ite < 2; ite++) {
                correct(H, g, t_c);
                if (merged)
                    for (int j = 0; j < n; j++) if (H.twin_of[b + j] == i) correct(H, b + j, t_c);
                gravity_pc(P, H, J, n, g, true);
                if (merged) {
                    for (int j = 0; j < n; j++) if (H.twin_of[b + j] == i) gravity_pc(P, H, J, n, b + j, true);
                    merge_acc_jerk(H, b, n, i);
                }
            }
            H.time[g] += t_c;
            if (merged)
                for (int j = 0; j < n; j++) if (H.twin_of[b + j] == i) H.time[b + j] += t_c;
            if (!flag_col) {
                double ad[3], jd[3], a2d[3], a3d[3], as[3], js[3], a2s[3], a3s[3];
                ld3(H.acc_d, g, ad); ld3(H.jerk_d, g, jd); ld3(H.a2_d, g, a2d); ld3(H.a3_d, g, a3d);
                ld3(H.acc_s, g, as); ld3(H.jerk_s, g, js); ld3(H.a2_s, g, a2s); ld3(H.a3_s, g, a3s);
                H.dt[g] = dt_block(P, H.time[g], H.dt[g], H.acc0[g], ad, jd, a2d, a3d, as, js, a2s, a3s);
                if (merged)
                    for (int j = 0; j < n; j++) if (H.twin_of[b + j] == i) H.dt[b + j] = H.dt[g];
            }
        }
        __threadfence_block();
        __syncthreads();
        if (flag_col) {
            collision_tail(P, H, b, n, ic, jc, ed, pairs);
            n_col++;
        }
        // getSystemTime
        double tm = 1.e10;
        for (int i = t; i < n; i += B)
            if (!(H.flags[b + i] & GPLUM_FLAG_DEAD)) tm = dmin(tm, H.time[b + i]);
        sys_time = COOP ? block_min_waves(tm, W->sys) : block_min(R, tm);
        loop++;
    }
    for (int i = t; i < n; i += B) { gravity_full(P, H, b, n, i); pairs += n - 1; }
    __threadfence_block();
    __syncthreads();
    if (n_col > 0) {
        for (int i = t; i < n; i += B)
            if (H.flags[b + i] & GPLUM_FLAG_MERGED) merge_acc_jerk(H, b, n, i);
    }
    __syncthreads();
}

// Pairs side by side for clusters of 3 to 8 particles (n (n - 1) <= 64): particle p
// is held by the m = n - 1 lanes p m .. p m + m - 1, each keeping the particle's whole
// state (computed alike), and lane p m + k evaluates p's pair with its k-th partner
// j = k + (k >= p).  One pair_term latency covers all pairs of a block step; every lane
// of the group then adds the group's m summands from the LDS, from zero in ascending j
// (the sum of gravity_loop_lane).  Called by the lanes of active particles only.
__device__ inline void side_force(const HP &P, int t, int base, int m, const double xo[3], const double vo[3],
                                  double rinv, const double xj[3], const double vj[3], double mj, double rinv_j,
                                  double *s_tm, double as[3], double js[3], double ad[3], double jd[3]) {
    double ta[3], tj[3];
    pair_term(P, xo, vo, rinv, xj[0], xj[1], xj[2], vj[0], vj[1], vj[2], mj, rinv_j, ta, tj, nullptr);
    for (int d = 0; d < 3; d++) { s_tm[6 * t + d] = ta[d]; s_tm[6 * t + 3 + d] = tj[d]; }
    star_acc_jerk(P, xo, vo, as, js, nullptr);
    wave_sync();
    for (int d = 0; d < 3; d++) { ad[d] = 0.; jd[d] = 0.; }
    for (int k = 0; k < m; k++) {
        const int L = base + k;
        for (int d = 0; d < 3; d++) { ad[d] += s_tm[6 * L + d]; jd[d] += s_tm[6 * L + 3 + d]; }
    }
    wave_sync();   // summands read before the next evaluation writes them
}

// wave_min2 over the particles' values when the lanes p m (p < n) hold them and the
// other lanes hold copies or 1e10 (m, n wave-uniform): up to four particles read
// their lanes directly.
__device__ inline void wave_min2_lead(double &u, double &v, int n, int m) {
    if (n <= 2) {
        u = dmin(readlane_f64(u, 0), readlane_f64(u, m));
        v = dmin(readlane_f64(v, 0), readlane_f64(v, m));
    } else if (n <= 4) {
        u = dmin(dmin(readlane_f64(u, 0), readlane_f64(u, m)), dmin(readlane_f64(u, 2 * m), readlane_f64(u, 3 * m)));
        v = dmin(dmin(readlane_f64(v, 0), readlane_f64(v, m)), dmin(readlane_f64(v, 2 * m), readlane_f64(v, 3 * m)));
    } else {
        wave_min2(u, v);
    }
}

// ------------------------------------------- clusters of 2 to 4 bodies: lane per term
//
// Lane L = 16 T + 4 q + c works on term T (the rows: T = 0 the Sun, T = k + 1 the pair
// with particle p's k-th partner in ascending index, j = k + (k >= p)), particle p = q
// (the quads of a row; for 2 bodies p = q & 1 and T = (L >> 4) & 1, so that every lane
// holds a copy) and component c (0, 1, 2 = x, y, z; c = 3 pads the quad and keeps
// zeros).  The Sun term and every pair term of every particle are evaluated by the same
// instructions, a vector operation acts on x, y and z at once, and a dot product is the
// quad's sum of its lanes' products in two DPP steps.  Every lane of particle p keeps
// p's state for its component (the same in all rows), so predictor and corrector need
// no exchange; after each evaluation every row gets each row's term of its (p, c) by
// cross-row lane swaps (v_permlane16/32_swap, no LDS) and adds them as star_acc_jerk
// and gravity_loop_lane do (the Sun row; the partner rows in ascending j from zero), and
// the Sun and planet time-step criteria likewise.  Same operations in the same order
// as the one-lane-per-particle path.

constexpr int DPP_ROW_ROR4 = 0x124;

// (s0 + s1) + s2 on every lane of a quad of lane values s0..s3, which is dot3's sum,
// when the padding value s3 is +0 and s2 is not -0 (squares), or s3 is -0 (any s2):
// s0 + s1 = s1 + s0 and the last addition commute, and s2 + s3 = s2.  All four lanes of
// the quad must be active.
__device__ inline double quad_sum(double s) {
    s = s + dpp_f64<DPP_QUAD_1032>(s);
    return s + dpp_f64<DPP_QUAD_2301>(s);
}
__device__ inline double quad_sq(double a) { return quad_sum(a * a); }
__device__ inline double quad_dot(double a, double b, bool pad) { return quad_sum(pad ? -0. : a * b); }

// Cross-row exchanges.  swap16: each lane gets the values of the even and the odd row
// of its pair of rows (0, 1) or (2, 3), same lane within the row; swap32: the values of
// its half's row at the same position in the lower and the upper half.  Combined, each
// row reads the four rows' values of its lane position.
__device__ inline double dbl_of(unsigned lo, unsigned hi) {
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ inline void swap16(double x, double &even, double &odd) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)u, (unsigned)u, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(u >> 32), (unsigned)(u >> 32), false, false);
    even = dbl_of(lo[0], hi[0]);
    odd = dbl_of(lo[1], hi[1]);
}
__device__ inline void swap32(double x, double &low, double &high) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)u, (unsigned)u, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(u >> 32), (unsigned)(u >> 32), false, false);
    low = dbl_of(lo[0], hi[0]);
    high = dbl_of(lo[1], hi[1]);
}
// the row values r[0..N-1] of the lane's position (N = 2: rows 0 and 1 of its pair)
template <int N>
__device__ inline void rows_of(double x, double (&r)[4]) {
    if (N == 2) {
        swap16(x, r[0], r[1]);
        r[2] = r[3] = 0.;
    } else {
        double lo, hi;
        swap32(x, lo, hi);
        swap16(lo, r[0], r[1]);
        swap16(hi, r[2], r[3]);
    }
}

// K(x) and dK/dr of cutfunc.h with selects (the polynomial evaluated everywhere).
__device__ inline double cutoff_K_sel(const Cutoff &c, double rij, double r_out_inv) {
    const double x = (c.gamma - rij * r_out_inv) * c.g_1_inv;
    const double x2 = x * x;
    const double p = (((-20. * x + 70.) * x - 84.) * x + 35.) * x2 * x2;
    return (x < 0.) ? 0. : ((x >= 1.) ? 1. : p);
}
__device__ inline double cutoff_dKdt_sel(const Cutoff &c, double rij, double r_out_inv, double alpha) {
    const double x = (c.gamma - rij * r_out_inv) * c.g_1_inv;
    const double x_1 = x - 1.;
    const double d = 140. * x * x * x * x_1 * x_1 * x_1 * r_out_inv * c.g_1_inv;
    const double dKdr = (x < 0. || x >= 1.) ? 0. : d;
    return alpha * rij * dKdr;
}

// The lane's term of a hard force evaluation, in one instruction stream for the Sun row
// (star_acc_jerk: xj = vj = -0, so that xj - x = -x exactly; mT = m_sun, eps2T =
// eps2_sun) and the partner rows (pair_term with r_out_inv = min of the two 1/r_out):
// the Sun's (1/s, then sqrt) and the pair's (sqrt, then 1/r) sequences are both
// evaluated and the row's own one selected; the Sun row takes K = 1 and dK/dt = -0,
// exact identities (1 K = K, alpha 1 = alpha, 3 alpha + (-0) = 3 alpha) that leave
// star_acc_jerk's acc = mj dr and jerk = mj (dv - 3 alpha dr).
__device__ inline void term_eval(const Cutoff &cut, bool sun, bool pad, double xi, double vi, double xj, double vj,
                                 double mT, double eps2T, double r_out_inv, double &ta, double &tj) {
    const double dr = xj - xi;
    const double dv = vj - vi;
    const double s = quad_sq(dr) + eps2T;
    const double a = 1. / s, b = sqrt(s);
    const double y1 = sun ? a : b;                 // Sun: 1/r^2; pair: r
    const double c = sqrt(y1), d = 1. / y1;
    const double rinv = sun ? c : d;
    const double r2inv = sun ? y1 : d * d;
    const double r3inv = r2inv * rinv;
    const double mj_r3 = mT * r3inv;
    const double alpha = quad_dot(dr, dv, pad) * r2inv;
    const double K = sun ? 1. : 1. - cutoff_K_sel(cut, y1, r_out_inv);
    const double dKdt = sun ? -0. : cutoff_dKdt_sel(cut, y1, r_out_inv, alpha);
    const double alpha_c = alpha * K;
    const double c2 = 3. * alpha_c + dKdt;
    ta = mj_r3 * K * dr;
    tj = mj_r3 * (K * dv - c2 * dr);
}

// Every row of particle p gets the Sun term and the sum of the partner terms of its
// component (rows 1..N-1, added in ascending j from zero).
template <int N>
__device__ inline void gather_terms(double ta, double tj, double &as, double &js, double &ad, double &jd) {
    double ra[4], rj[4];
    rows_of<N>(ta, ra);
    rows_of<N>(tj, rj);
    as = ra[0];
    js = rj[0];
    ad = 0.;
    jd = 0.;
#pragma unroll
    for (int k = 1; k < N; k++) {
        ad += ra[k];
        jd += rj[k];
    }
}

// predict_regs and correct_regs for one component (same expressions).
__device__ inline void predict_one(double Dt, double pos, double vel, double as, double js, double ad, double jd,
                                   double &xp, double &vp) {
    const double acc_sd = as + ad;
    const double jerk_sd = js + jd;
    xp = pos + Dt * (vel + Dt * (SECONDORDER * acc_sd + THIRDORDER * Dt * jerk_sd));
    vp = vel + Dt * (acc_sd + Dt * SECONDORDER * jerk_sd);
}
__device__ inline void correct_one(double Dt, double Dt2, double Dt_inv2, double Dt_inv3, double xp, double vp,
                                   double as, double a0s, double j0s, double js, double ad, double a0d, double j0d,
                                   double jd, double &pos, double &vel, double &a2s, double &a3s, double &a2d,
                                   double &a3d) {
    const double Am_s = as - a0s;
    const double J0_s = Dt * j0s;
    const double J1_s = Dt * js;
    const double A2_s = 3. * Am_s - (J1_s + 2. * J0_s);
    const double A3_s = -2. * Am_s + (J1_s + J0_s);
    a2s = 2. * Dt_inv2 * (A2_s + 3. * A3_s);
    a3s = 6. * Dt_inv3 * A3_s;
    const double Am_d = ad - a0d;
    const double J0_d = j0d * Dt;
    const double J1_d = jd * Dt;
    const double A2_d = 3. * Am_d - (J1_d + 2. * J0_d);
    const double A3_d = -2. * Am_d + (J1_d + J0_d);
    a2d = 2. * (A2_d + 3. * A3_d) * Dt_inv2;
    a3d = 6. * A3_d * Dt_inv3;
    const double A2_sd = A2_s + A2_d;
    const double A3_sd = A3_s + A3_d;
    pos = xp + ONE_SIXTIETH * Dt2 * (5. * A2_sd + 3. * ALPHA_POS * A3_sd);
    vel = vp + ONE_TWELFTH * Dt * (4. * A2_sd + 3. * A3_sd);
}

// calc_dt4_fast and calc_dt4 of the lane's term (the row's vectors, dots over the quad).
__device__ inline double dt4_fast_quad(double eta, double alpha2, double acc0, double a, double j, double s,
                                       double c) {
    const double Acc2 = quad_sq(a) + alpha2 * acc0 * acc0;
    const double Jerk2 = quad_sq(j);
    const double Snap2 = quad_sq(s);
    const double Crac2 = quad_sq(c);
    const double as2 = Acc2 * Snap2, jc2 = Jerk2 * Crac2;
    const bool ok = (as2 >= DBL_MIN || (Acc2 == 0. || Snap2 == 0.)) && (jc2 >= DBL_MIN || Crac2 == 0.);
    const double num = __builtin_amdgcn_sqrt(as2) + Jerk2;
    const double den = __builtin_amdgcn_sqrt(jc2) + Snap2;
    const double q = ok ? eta * __builtin_amdgcn_sqrt(num * __builtin_amdgcn_rcp(den)) : __builtin_nan("");
    return (Jerk2 > 0.) ? q : DBL_MAX;
}
__device__ inline double dt4_quad(double eta, double alpha2, double acc0, double a, double j, double s, double c) {
    const double Acc = sqrt(quad_sq(a) + alpha2 * acc0 * acc0);
    const double Jerk2 = quad_sq(j);
    const double Jerk = sqrt(Jerk2);
    const double Snap2 = quad_sq(s);
    const double Snap = sqrt(Snap2);
    const double Crac = sqrt(quad_sq(c));
    return (Jerk > 0.) ? eta * sqrt((Acc * Snap + Jerk2) / (Jerk * Crac + Snap2)) : DBL_MAX;
}

// Physics constants of the lane-per-term path, in registers (the kernel keeps HP in LDS).
struct TermConsts {
    Cutoff cut;
    double dt_tree, dt_min, alpha2;
};

// the Sun row's and the first partner row's values of the lane's position
template <int N>
__device__ inline void sun_planet_of(double x, double &f_sun, double &f_planet) {
    if (N == 2) {
        swap16(x, f_sun, f_planet);
    } else {
        double lo, hi;
        swap32(x, lo, hi);
        swap16(lo, f_sun, f_planet);
    }
}

// dt_block with the Sun criterion from the Sun row and the planet criterion from the
// partner rows (each lane evaluates its own row's, reads the other by a lane swap); the
// exact criteria when either fast one is not finite or dt_1 lies near the step grid.
template <int N>
__device__ inline double dt_block_rows(const TermConsts &C, double time, double dt, double etaT, double acc0T,
                                       double a, double j, double s, double c) {
    double dt_next = dmin(2. * dt, 0.5 * C.dt_tree);
    double rem = fmod_block(time, dt_next);
    for (int it = 0; rem != 0.0 && it < HALVE_CAP; it++) { dt_next *= 0.5; rem = fmod_block(time, dt_next); }
    double f_sun, f_planet;
    sun_planet_of<N>(dt4_fast_quad(etaT, C.alpha2, acc0T, a, j, s, c), f_sun, f_planet);
    double dt_1 = dmin(f_planet, f_sun);
    const bool exact = !(f_planet <= DBL_MAX && f_sun <= DBL_MAX) || !clear_of_grid(dt_1, dt_next);
    if (__ballot(exact)) {
        double e_sun, e_planet;
        sun_planet_of<N>(dt4_quad(etaT, C.alpha2, acc0T, a, j, s, c), e_sun, e_planet);
        if (exact) dt_1 = dmin(e_planet, e_sun);
    }
    for (int it = 0; dt_1 < dt_next && 0.5 * dt < dt_next && it < HALVE_CAP; it++) dt_next *= 0.5;
    if (dt_next < 2. * C.dt_min) dt_next = C.dt_min;
    return (rem != 0.0) ? C.dt_min : dt_next;
}

// Block steps of a cluster of N (2 to 4) particles with distinct ids from the start of
// the soft step (the layout above; one wave).  Returns false when a block step has a
// collision: the state of the start of that block step is then in H (flags and twins
// cleared) for integrate_from.  Else the final calcGravity with potentials is done and
// written to H.  s_m, s_ri, s_id, s_fr hold the cluster's masses, 1/r_out, ids and
// f r_planet; s_x, s_v and s_tm are free.
template <int N>
__device__ bool integrate_terms(const HP &P, HardView &H, const JData &J, double *s_x, double *s_v, double *s_m,
                                double *s_ri, int64_t *s_id, double *s_fr, double *s_tm, int b, double &sys_time,
                                long long &loop, long long &pairs) {
    constexpr int m = N - 1;
    const int L = threadIdx.x;
    const int T = (N == 2) ? ((L >> 4) & 1) : (L >> 4);
    const int p = (N == 2) ? ((L >> 2) & 1) : ((L >> 2) & 3);
    const int c = L & 3;
    // lanes of a particle: every row takes part in the evaluations, so that the cross-row
    // swaps read computed values (rows T >= N, 3 bodies, evaluate a don't-care term)
    const bool own = p < N;
    const bool row = own && T < N;                  // lanes of a term of a particle
    const bool first = (N == 2) ? L < 32 && ((L >> 2) & 3) < 2 : true;   // not a copy
    const bool sun = T == 0, pad = c == 3;
    const int j = (T >= 1) ? (T - 1) + ((T - 1) >= p ? 1 : 0) : p;   // the row's partner
    const int jc = (j < N) ? j : 0, pc = own ? p : 0;
    const int g = b + pc;
    const TermConsts C{P.cut, P.dt_tree, P.dt_min, P.alpha2};
    const double mT = sun ? P.m_sun : s_m[jc];
    const double eps2T = sun ? P.eps2_sun : P.eps2;
    const double r_out_inv = dmin(s_ri[pc], s_ri[jc]);
    const double etaT = sun ? P.eta_sun : P.eta;
    const double r2c = s_fr[pc] + s_fr[jc];
    const double col_thr = COL_FILTER * r2c * r2c;
    const bool col_row = P.collision && row && !sun;
    const int src = 4 * jc + c;                     // a lane of the partner's component
    const double time_end = P.dt_tree;
    double x = 0., v = 0., time = 0., dt = 0., acc0 = 0.;
    if (own && !pad) {
        x = H.pos[3 * g + c];
        v = H.vel[3 * g + c];
    }
    if (own) {
        time = H.time[g];
        dt = H.dt[g];
        acc0 = H.acc0[g];
    }
    const double acc0T = sun ? 0. : acc0;
    const bool lead = row && sun && c == 0 && first;
    double as = 0., js = 0., ad = 0., jd = 0., a2s = 0., a3s = 0., a2d = 0., a3d = 0.;
    {
        // initial calcGravity (hermite.h:259-270) at the current state
        double xj = __shfl(x, src), vj = __shfl(v, src);
        if (sun) { xj = -0.; vj = -0.; }
        if (own) {
            double ta, tj;
            term_eval(C.cut, sun, pad, x, v, xj, vj, mT, eps2T, r_out_inv, ta, tj);
            gather_terms<N>(ta, tj, as, js, ad, jd);
        }
    }
    double tn = own ? time + dt : 1.e10;
    tn = dmin(tn, dpp_f64<DPP_ROW_ROR4>(tn));
    tn = dmin(tn, dpp_f64<DPP_ROW_ROR8>(tn));
    double t_s = readlane_f64(tn, 0);
    bool handover = false;
    while (sys_time < time_end && loop < BLOCK_STEP_CAP) {
        const double Dt = t_s - time;
        const double a0s = as, j0s = js, a0d = ad, j0d = jd;
        double xp = 0., vp = 0.;
        predict_one(Dt, x, v, as, js, ad, jd, xp, vp);
        double xj, vj;
        if (N == 2) {
            // the partner of quad q is quad q ^ 1, which row_ror:4 brings (quads 2, 3 are
            // copies of 0, 1)
            xj = dpp_f64<DPP_ROW_ROR4>(xp);
            vj = dpp_f64<DPP_ROW_ROR4>(vp);
        } else {
            xj = __shfl(xp, src);
            vj = __shfl(vp, src);
        }
        if (sun) { xj = -0.; vj = -0.; }
        if (P.collision) {
            // collisionDetermination: a pair with R = |dx| / (f_i R_i + f_j R_j) < 1
            const double dx = xp - xj;
            const double d2 = quad_sq(dx);
            const bool near = col_row && d2 < col_thr;
            if (__ballot(near)) {
                const bool hit = near && sqrt(d2) / r2c < 1.;
                if (__ballot(hit)) { handover = true; break; }
            }
        }
        if (own && time + dt == t_s) {
            if (lead) pairs += 3 * (long long)m;
            const double Dt2 = Dt * Dt;
            const double Dt_inv = recip_block(Dt);
            const double Dt_inv2 = Dt_inv * Dt_inv;
            const double Dt_inv3 = Dt_inv2 * Dt_inv;
            double ta, tj;
            term_eval(C.cut, sun, pad, xp, vp, xj, vj, mT, eps2T, r_out_inv, ta, tj);
            gather_terms<N>(ta, tj, as, js, ad, jd);
            for (int ite = 0; ite < 2; ite++) {
                correct_one(Dt, Dt2, Dt_inv2, Dt_inv3, xp, vp, as, a0s, j0s, js, ad, a0d, j0d, jd, x, v, a2s, a3s,
                            a2d, a3d);
                term_eval(C.cut, sun, pad, x, v, xj, vj, mT, eps2T, r_out_inv, ta, tj);
                gather_terms<N>(ta, tj, as, js, ad, jd);
            }
            time += Dt;
            dt = dt_block_rows<N>(C, time, dt, etaT, acc0T, sun ? as : ad, sun ? js : jd, sun ? a2s : a2d,
                                  sun ? a3s : a3d);
        }
        double tm = own ? time : 1.e10;
        tn = own ? time + dt : 1.e10;
        tm = dmin(tm, dpp_f64<DPP_ROW_ROR4>(tm));
        tn = dmin(tn, dpp_f64<DPP_ROW_ROR4>(tn));
        tm = dmin(tm, dpp_f64<DPP_ROW_ROR8>(tm));
        tn = dmin(tn, dpp_f64<DPP_ROW_ROR8>(tn));
        sys_time = readlane_f64(tm, 0);
        t_s = readlane_f64(tn, 0);
        loop++;
    }
    if (handover) {
        // the state at the start of this block step, for integrate_from
        if (row && sun && !pad && first) {
            H.pos[3 * g + c] = x; H.vel[3 * g + c] = v;
            H.acc_s[3 * g + c] = as; H.jerk_s[3 * g + c] = js;
            H.acc_d[3 * g + c] = ad; H.jerk_d[3 * g + c] = jd;
        }
        if (lead) {
            H.time[g] = time;
            H.dt[g] = dt;
            H.flags[g] &= ~(GPLUM_FLAG_DEAD | GPLUM_FLAG_MERGED);
            H.twin_of[g] = -1;
        }
        return false;
    }
    // final calcGravity with potentials (hermite.h:466-470), one lane per particle
    __syncthreads();
    if (row && sun && !pad && first) {
        s_x[3 * p + c] = x;
        s_v[3 * p + c] = v;
    }
    if (lead) {
        s_tm[2 * p] = time;
        s_tm[2 * p + 1] = dt;
    }
    __syncthreads();
    if (L < N) {
        const int gi = b + L;
        double xi[3], vi[3], fas[3], fjs[3], fad[3], fjd[3], phs = 0., phd = 0.;
        ld3(s_x, L, xi);
        ld3(s_v, L, vi);
        star_acc_jerk(P, xi, vi, fas, fjs, &phs);
        gravity_loop_lane<false>(P, J, N, L, s_id[L], xi, vi, s_ri[L], fad, fjd, &phd);
        pairs += m;
        st3(H.pos, gi, xi); st3(H.vel, gi, vi);
        st3(H.acc_s, gi, fas); st3(H.jerk_s, gi, fjs);
        st3(H.acc_d, gi, fad); st3(H.jerk_d, gi, fjd);
        H.time[gi] = s_tm[2 * L];
        H.dt[gi] = s_tm[2 * L + 1];
        H.phi_s[gi] = phs;
        H.phi_d[gi] = phd;
        H.flags[gi] &= ~(GPLUM_FLAG_DEAD | GPLUM_FLAG_MERGED);
        H.twin_of[gi] = -1;
    }
    __syncthreads();
    return true;
}

// Clusters that fit one wave: each particle's whole state stays in registers (one
// lane per particle, or a group of lanes for 3 to 8 particles, see side_force) and
// only the predicted j-data (and f r_planet) goes through LDS, so a block step costs a
// few LDS round trips and wave shuffles instead of global-memory round trips.  Same
// operations in the same order as integrate_from.  At the first block step with a
// collision the state is written back to H and integrate_from takes over from the
// start of that block step (it predicts the same positions and finds the same pair),
// so merging and twins need no register path.
__device__ void integrate_regs(const HP &P, HardView &H, const JData &J, Reduce &R, double *s_x, double *s_v,
                               double *s_m, double *s_ri, int64_t *s_id, double *s_fr, double *s_tm, int b,
                               int n, double &sys_time, long long &loop, double &ed, int &n_col,
                               long long &pairs) {
    n = __builtin_amdgcn_readfirstlane(n);   // wave-uniform: scalar branches on the size
    const int t = threadIdx.x;
    const bool side = n >= 3 && n * (n - 1) <= 64;
    const int m = side ? n - 1 : 1;            // lanes per particle
    const int p = t / m;                       // this lane's particle
    const bool own = p < n;
    const bool lead = own && t == p * m;       // the lane that writes p's shared state
    const int g = b + p;
    double x[3] = {0., 0., 0.}, v[3] = {0., 0., 0.}, as[3], js[3], ad[3], jd[3];
    double time = 0., dt = 0., acc0 = 0., rinv = 0., fr = 0.;
    int64_t id = -1;
    if (own) {
        ld3(H.pos, g, x);
        ld3(H.vel, g, v);
        time = H.time[g];
        dt = H.dt[g];
        acc0 = H.acc0[g];
        rinv = H.rinv[g];
        id = H.id[g];
        fr = H.f[g] * H.r_planet[g];
        if (lead) {
            s_m[p] = H.mass[g];
            s_ri[p] = rinv;
            s_id[p] = id;
            s_fr[p] = fr;
            st3(s_x, p, x);
            st3(s_v, p, v);
        }
    }
    __syncthreads();
    // The clusters the reference forms have distinct ids, and the partner loops below
    // leave the id test out; a cluster with a repeated id takes the global-memory path
    // from the start (the non-register kernels' preamble, then integrate_from).
    // (One call site of integrate_from serves both hand-overs: a second inlined copy
    // would double the kernel's register allocation.)
    bool rep = false;
    if (own)
        for (int j = 0; j < n; j++) rep = rep || (j != p && s_id[j] == id);
    const bool from_start = __ballot(rep) != 0;
    // 2 to 4 bodies: the lane-per-term path (a collision hands over to integrate_from below)
    const bool terms = !from_start && n <= P.terms_max;
    bool terms_handover = false;
    if (terms) {
        const bool done =
            (n == 2) ? integrate_terms<2>(P, H, J, s_x, s_v, s_m, s_ri, s_id, s_fr, s_tm, b, sys_time, loop, pairs)
            : (n == 3) ? integrate_terms<3>(P, H, J, s_x, s_v, s_m, s_ri, s_id, s_fr, s_tm, b, sys_time, loop, pairs)
                       : integrate_terms<4>(P, H, J, s_x, s_v, s_m, s_ri, s_id, s_fr, s_tm, b, sys_time, loop, pairs);
        if (done) return;
        terms_handover = true;
    }
    // side_force: this lane's partner and its constants
    int jp = 0;
    double m_jp = 0., ri_jp = 0.;
    if (side && own) {
        const int k = t - p * m;
        jp = k + (k >= p ? 1 : 0);
        m_jp = s_m[jp];
        ri_jp = s_ri[jp];
    }
    if (from_start) {
        if (lead) {
            gravity_full(P, H, b, n, p);
            H.flags[g] &= ~(GPLUM_FLAG_DEAD | GPLUM_FLAG_MERGED);
            H.twin_of[g] = -1;
        }
    } else if (own && !terms) {
        // initial calcGravity (hermite.h:259-270)
        star_acc_jerk(P, x, v, as, js, nullptr);
        gravity_loop_lane<false>(P, J, n, p, id, x, v, rinv, ad, jd, nullptr);
    }
    const double time_end = P.dt_tree;
    double a0s[3], j0s[3], a0d[3], j0d[3], xp[3], vp[3], a2s[3], a3s[3], a2d[3], a3d[3];
    // t_s: the next block time, min(time + dt) (makeActiveList); computed together with
    // the system time at the end of each step
    double t_s = wave_min(own ? time + dt : 1.e10);
    bool handover = false;
    while (!from_start && !terms && sys_time < time_end && loop < BLOCK_STEP_CAP) {
        __syncthreads();   // the previous step's j-data reads are done
        if (own) {
            predict_regs(t_s - time, x, v, as, js, ad, jd, a0s, j0s, a0d, j0d, xp, vp);
            if (lead) {
                st3(s_x, p, xp);
                st3(s_v, p, vp);
            }
        }
        __syncthreads();
        if (P.collision) {
            double rmin = 1.;
            int bi = 0x7fffffff, bj = -1;
            if (own) {
                for (int k = 0; k < n - 1; k++) {
                    const int j = k + (k >= p ? 1 : 0);   // the partners, ascending
                    const double dx = xp[0] - s_x[3 * j];
                    const double dy = xp[1] - s_x[3 * j + 1];
                    const double dz = xp[2] - s_x[3 * j + 2];
                    const double d2 = dot3(dx, dy, dz, dx, dy, dz);
                    const double r2 = fr + s_fr[j];
                    if (!(d2 < COL_FILTER * r2 * r2)) continue;   // R >= 1: never the minimum
                    const double Rr = sqrt(d2) / r2;
                    if (Rr < rmin) { rmin = Rr; bi = p; bj = j; }
                }
            }
            // the argmin only when some pair is below 1 (rare)
            if (__ballot(rmin < 1.)) wave_argmin(rmin, bi, bj);
            if (__ballot(rmin < 1.)) { handover = true; break; }
        }
        if (own && time + dt == t_s) {
            if (lead) pairs += 3 * (long long)(n - 1);
            const double t_c = t_s - time;
            if (side) {
                // the partner's predicted state, once for the three evaluations
                double xj[3], vj[3];
                ld3(s_x, jp, xj);
                ld3(s_v, jp, vj);
                side_force(P, t, p * m, m, xp, vp, rinv, xj, vj, m_jp, ri_jp, s_tm, as, js, ad, jd);
                for (int ite = 0; ite < 2; ite++) {
                    correct_regs(t_c, xp, vp, as, a0s, j0s, js, ad, a0d, j0d, jd, x, v, a2s, a3s, a2d, a3d);
                    side_force(P, t, p * m, m, x, v, rinv, xj, vj, m_jp, ri_jp, s_tm, as, js, ad, jd);
                }
            } else if (n == 2) {
                // the partner's predicted state is read once for the three evaluations
                JReg pj[1];
                load_jreg(J, p == 0 ? 1 : 0, pj[0]);
                force_jreg<1>(P, pj, xp, vp, rinv, as, js, ad, jd);
                for (int ite = 0; ite < 2; ite++) {
                    correct_regs(t_c, xp, vp, as, a0s, j0s, js, ad, a0d, j0d, jd, x, v, a2s, a3s, a2d, a3d);
                    force_jreg<1>(P, pj, x, v, rinv, as, js, ad, jd);
                }
            } else {
                star_acc_jerk(P, xp, vp, as, js, nullptr);
                gravity_loop_lane<false>(P, J, n, p, id, xp, vp, rinv, ad, jd, nullptr);
                for (int ite = 0; ite < 2; ite++) {
                    correct_regs(t_c, xp, vp, as, a0s, j0s, js, ad, a0d, j0d, jd, x, v, a2s, a3s, a2d, a3d);
                    star_acc_jerk(P, x, v, as, js, nullptr);
                    gravity_loop_lane<false>(P, J, n, p, id, x, v, rinv, ad, jd, nullptr);
                }
            }
            time += t_c;
            dt = dt_block(P, time, dt, acc0, ad, jd, a2d, a3d, as, js, a2s, a3s);
        }
        sys_time = own ? time : 1.e10;
        t_s = own ? time + dt : 1.e10;
        wave_min2_lead(sys_time, t_s, n, m);
        loop++;
    }
    if (handover || from_start || terms_handover) {
        // a collision in this block step: hand the cluster to the global-memory path
        // at the start of the step (kept out of the loop, which then holds less state)
        if (handover && lead) {
            st3(H.pos, g, x); st3(H.vel, g, v);
            st3(H.acc_s, g, as); st3(H.jerk_s, g, js);
            st3(H.acc_d, g, ad); st3(H.jerk_d, g, jd);
            H.time[g] = time;
            H.dt[g] = dt;
            H.flags[g] &= ~(GPLUM_FLAG_DEAD | GPLUM_FLAG_MERGED);
            H.twin_of[g] = -1;
        }
        __threadfence_block();
        __syncthreads();
        integrate_from<true, false>(P, H, J, R, s_x, s_v, s_id, s_fr, b, n, sys_time, loop, ed, n_col, pairs);
        return;
    }
    // final calcGravity with potentials (hermite.h:466-470), current positions
    __syncthreads();
    if (lead) {
        st3(s_x, p, x);
        st3(s_v, p, v);
    }
    __syncthreads();
    if (own) {
        double phs = 0., phd = 0.;
        star_acc_jerk(P, x, v, as, js, &phs);
        gravity_loop_lane<false>(P, J, n, p, id, x, v, rinv, ad, jd, &phd);
        if (lead) {
            pairs += n - 1;
            st3(H.pos, g, x); st3(H.vel, g, v);
            st3(H.acc_s, g, as); st3(H.jerk_s, g, js);
            st3(H.acc_d, g, ad); st3(H.jerk_d, g, jd);
            H.time[g] = time;
            H.dt[g] = dt;
            H.phi_s[g] = phs;
            H.phi_d[g] = phd;
            H.flags[g] &= ~(GPLUM_FLAG_DEAD | GPLUM_FLAG_MERGED);
            H.twin_of[g] = -1;
        }
    }
    __syncthreads();
}

// Three instantiations, one per cluster-size class (each skips the others):
// REG: n <= HARD_SEQ_MAX, particle state in registers (integrate_regs);
// LDSJ && !REG: HARD_SEQ_MAX < n <= HERMITE_LDS, j-data in LDS (integrate_from);
// !LDSJ: n > HERMITE_LDS, j-data in global memory.
// The REG kernel runs one wave per workgroup; the other two run HERMITE_BIG_BLOCK
// threads, whose waves share the active particles of a block step.
// Separate kernels keep the register-heavy small-cluster path from lowering the
// occupancy of the others.
template <bool LDSJ, bool REG>
__device__ __forceinline__ void hermite_clusters(HP P_arg, HardView H,
                                                           const int32_t *__restrict__ cl_off,
                                                           const int32_t *__restrict__ order, int ncl,
                                                           int *__restrict__ next, double *__restrict__ edisp,
                                                           int *__restrict__ ncol, long long *__restrict__ nblock,
                                                           long long *__restrict__ npair, int persistent,
                                                           int lds_max, int giant_min) {
    __shared__ double s_x[3 * HERMITE_LDS], s_v[3 * HERMITE_LDS], s_m[HERMITE_LDS], s_ri[HERMITE_LDS];
    __shared__ int64_t s_id[HERMITE_LDS];
    __shared__ double s_fr[HERMITE_LDS];
    __shared__ Reduce R;
    __shared__ int s_cl;
    __shared__ unsigned long long s_pairs;
    // the register path reads the physics constants from LDS: as kernel arguments they
    // would occupy scalar registers for the whole kernel (and be spilled and reloaded
    // in its loop); the other paths keep them as arguments
    __shared__ HP s_P;
    // register path, 3 to 8 particles: pair summands (side_force)
    __shared__ double s_tm[REG ? 6 * 64 : 1];
    // multi-wave kernels: reduction slots and the active list, and each particle's
    // acc/jerk at the start of the block step (12 per particle) for the lane groups
    __shared__ __align__(8) unsigned char s_ws[REG ? 8 : sizeof(WaveSlots)];
    __shared__ double s_a0[(REG || !LDSJ) ? 1 : 12 * HERMITE_LDS];
    const int t = threadIdx.x;
    const int B = blockDim.x;
    if (REG) {
        if (t == 0) s_P = P_arg;
        __syncthreads();
    }
    const HP &P = REG ? s_P : P_arg;
    for (int iter = 0;; iter++) {
        if (persistent) {
            if (t == 0) s_cl = atomicAdd(next, 1);
            __syncthreads();
        } else {
            if (t == 0) s_cl = (iter == 0) ? (int)blockIdx.x : ncl;
            __syncthreads();
        }
        const int q = s_cl;
        __syncthreads();
        if (q >= ncl) break;
        const int c = order ? order[q] : q;
        const int b = cl_off[c];
        const int n = cl_off[c + 1] - b;
        if (REG && n > HARD_SEQ_MAX) continue;
        if (LDSJ && !REG && (n <= HARD_SEQ_MAX || n > lds_max || n > giant_min)) continue;
        if (!LDSJ && (n <= lds_max || n > giant_min)) continue;
        JData J;
        if (LDSJ) {
            J = JData{s_x, s_v, s_m, s_ri, s_id, 3};
            for (int i = t; i < n; i += B) { s_m[i] = H.mass[b + i]; s_ri[i] = H.rinv[b + i]; }
        } else {
            J = JData{H.xp + 3 * b, H.vp + 3 * b, H.mass + b, H.rinv + b, H.id + b, 3};
        }
        // pair evaluations of this cluster (hard_pair calls), for the roofline
        long long pairs = (long long)(n > 0 ? n - 1 : 0) * ((t < n) ? (n - 1 - t) / B + 1 : 0);
        if (t == 0) s_pairs = 0ull;
        double ed = 0.;
        int n_col = 0;
        long long loop = 0;
        double sys_time = 0.;
        if (REG) {
            integrate_regs(P, H, J, R, s_x, s_v, s_m, s_ri, s_id, s_fr, s_tm, b, n, sys_time, loop, ed, n_col,
                           pairs);
        } else {
            // initial calcGravity for all, flags cleared (hermite.h:259-270)
            for (int i = t; i < n; i += B) {
                gravity_full(P, H, b, n, i);
                H.flags[b + i] &= ~(GPLUM_FLAG_DEAD | GPLUM_FLAG_MERGED);
                H.twin_of[b + i] = -1;
            }
            __syncthreads();
            integrate_from<LDSJ, true>(P, H, J, R, s_x, s_v, s_id, s_fr, b, n, sys_time, loop, ed, n_col, pairs,
                                       reinterpret_cast<WaveSlots *>(s_ws), s_a0);
        }
        atomicAdd(&s_pairs, (unsigned long long)pairs);
        __syncthreads();
        const double time_end = P.dt_tree;
        if (t == 0) {
            edisp[c] = ed;
            ncol[c] = n_col;
            nblock[c] = (sys_time < time_end) ? -1 : loop;   // -1: block-step cap reached
            npair[c] = (long long)s_pairs;
        }
        __syncthreads();
    }
}

// The register path: one wave per workgroup, all the registers a lane may have.
template <bool LDSJ, bool REG>
__global__ __launch_bounds__(HERMITE_BLOCK) void k_hermite(HP P, HardView H, const int32_t *__restrict__ cl_off,
                                                           const int32_t *__restrict__ order, int ncl,
                                                           int *__restrict__ next, double *__restrict__ edisp,
                                                           int *__restrict__ ncol, long long *__restrict__ nblock,
                                                           long long *__restrict__ npair, int persistent, int lds_max,
                                                           int giant_min) {
    hermite_clusters<LDSJ, REG>(P, H, cl_off, order, ncl, next, edisp, ncol, nblock, npair, persistent, lds_max,
                                giant_min);
}

// The larger clusters' kernels: HERMITE_BIG_BLOCK threads, two waves per SIMD (256
// registers per lane).
template <bool LDSJ>
__global__ __launch_bounds__(HERMITE_BIG_BLOCK) __attribute__((amdgpu_waves_per_eu(2))) void k_hermite_multi(
    HP P, HardView H, const int32_t *__restrict__ cl_off, const int32_t *__restrict__ order, int ncl,
    int *__restrict__ next, double *__restrict__ edisp, int *__restrict__ ncol, long long *__restrict__ nblock,
    long long *__restrict__ npair, int persistent, int lds_max, int giant_min) {
    hermite_clusters<LDSJ, false>(P, H, cl_off, order, ncl, next, edisp, ncol, nblock, npair, persistent, lds_max,
                                  giant_min);
}

// ------------------------------------------------------------- isolated

// timeIntegrateKepler_isolated (hermite.h:787-815) / timeIntegrate_isolated
// (hermite.h:734-785) chosen as in hard.h:803-811.
struct IsoView {
    double *pos, *vel, *acc_d, *jerk_d, *acc_s, *jerk_s, *phi_d, *phi_s, *time, *dt, *acc0;
    const int32_t *number;
    int32_t *n_cluster;
};

__device__ double solve_kepler(double l, double ecc) {
    const double ecc2 = ecc * ecc, ecc3 = ecc2 * ecc, ecc4 = ecc2 * ecc2, ecc5 = ecc3 * ecc2, ecc6 = ecc3 * ecc3;
    double u = l + (ecc - ecc3 / 8. + ecc5 / 192.) * sin(l) + (ecc2 / 2. - ecc4 / 6. + ecc6 / 48.) * sin(2. * l) +
               (3. * ecc3 / 8. - 27. * ecc5 / 128.) * sin(3. * l) + (ecc4 / 3. - 4. * ecc6 / 15.) * sin(4. * l) +
               125. * ecc5 / 384. * sin(5. * l) + 27. * ecc6 / 80. * sin(6. * l);
    if (fabs((u - ecc * sin(u)) - l) > 1.e-15) {
        double u0;
        int loop = 0;
        do {
            u0 = u;
            const double s = sin(u0), c = cos(u0);
            u = u0 - ((u0 - ecc * s - l) / (1. - ecc * c));
            loop++;
        } while (fabs(u - u0) > 1.e-15 && loop < 10);
    }
    return u;
}

__device__ inline double eccentricity(const HP &P, const double x[3], const double v[3]) {
    const double r = sqrt(dot3(x[0], x[1], x[2], x[0], x[1], x[2]));
    const double rv = dot3(x[0], x[1], x[2], v[0], v[1], v[2]);
    const double ax = 1.0 / (2.0 / r - dot3(v[0], v[1], v[2], v[0], v[1], v[2]) / P.m_sun);
    const double ecccosu = 1. - r / ax;
    const double eccsinu2 = rv * rv / (P.m_sun * ax);
    return sqrt(ecccosu * ecccosu + eccsinu2);
}

__device__ void kepler_drift(const HP &P, double x[3], double v[3], double dtime) {
    const double mu = P.m_sun;
    const double r = sqrt(dot3(x[0], x[1], x[2], x[0], x[1], x[2]));
    const double rinv = 1. / r;
    const double v2 = dot3(v[0], v[1], v[2], v[0], v[1], v[2]);
    const double rv = dot3(x[0], x[1], x[2], v[0], v[1], v[2]);
    const double ax = 1.0 / (2.0 * rinv - v2 / mu);
    const double ecccosu = 1. - r / ax;
    const double eccsinu = rv / sqrt(mu * ax);
    const double ecc = sqrt(ecccosu * ecccosu + eccsinu * eccsinu);
    const double nn = sqrt(mu / (ax * ax * ax));
    double u, cosu, sinu;
    if (ecc != 0) { u = atan2(eccsinu, ecccosu); cosu = ecccosu / ecc; sinu = eccsinu / ecc; }
    else { u = 0.; cosu = 1.; sinu = 0.; }
    const double aninv = sqrt(ax / mu);
    const double ecc_sq = sqrt(1. - ecc * ecc);
    double Pv[3], Qv[3];
    for (int k = 0; k < 3; k++) {
        Pv[k] = rinv * cosu * x[k] - aninv * sinu * v[k];
        Qv[k] = (rinv * sinu * x[k] + aninv * (cosu - ecc) * v[k]) / ecc_sq;
    }
    double l = u - ecc * sin(u);
    l += nn * dtime;
    u = solve_kepler(l, ecc);
    const double c2 = cos(u), s2 = sin(u);
    for (int k = 0; k < 3; k++) x[k] = ax * ((c2 - ecc) * Pv[k] + ecc_sq * s2 * Qv[k]);
    const double ri2 = sqrt(1. / dot3(x[0], x[1], x[2], x[0], x[1], x[2]));
    for (int k = 0; k < 3; k++) v[k] = ax * ax * nn * ri2 * (-s2 * Pv[k] + ecc_sq * c2 * Qv[k]);
}

__global__ void k_isolated(HP P, IsoView V, int n, double time_start, double time_end) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || V.number[i] != 0) return;
    double x[3], v[3], as[3], js[3], phs;
    ld3(V.pos, i, x);
    ld3(V.vel, i, v);
    const double zero[3] = {0., 0., 0.};
    if (eccentricity(P, x, v) < 0.8 && P.eps2_sun == 0.) {
        kepler_drift(P, x, v, time_end - time_start);
        V.time[i] += (time_end - time_start);
        st3(V.pos, i, x);
        st3(V.vel, i, v);
        V.phi_d[i] = 0.;
        st3(V.acc_d, i, zero);
        st3(V.jerk_d, i, zero);
        star_acc_jerk(P, x, v, as, js, &phs);
        st3(V.acc_s, i, as);
        st3(V.jerk_s, i, js);
        V.phi_s[i] = phs;
        V.dt[i] = dt_initial(P, V.time[i], V.dt[i], V.acc0[i], zero, zero, as, js);
    } else {
        // Sun-only Hermite in local time (FPHard: time_c = time, time = 0)
        const double time_c = V.time[i];
        double time = 0., dt = V.dt[i];
        star_acc_jerk(P, x, v, as, js, nullptr);
        double ad[3] = {0., 0., 0.}, jd[3] = {0., 0., 0.};
        ld3(V.acc_d, i, ad);
        ld3(V.jerk_d, i, jd);
        double a2s[3], a3s[3], a2d[3], a3d[3];
        for (long long it = 0; time < P.dt_tree && it < BLOCK_STEP_CAP; it++) {
            const double Dt = dt;
            double xp[3], vp[3], a0s[3], j0s[3], a0d[3], j0d[3];
            for (int k = 0; k < 3; k++) {
                a0s[k] = as[k]; j0s[k] = js[k]; a0d[k] = ad[k]; j0d[k] = jd[k];
                const double acc_sd = as[k] + ad[k], jerk_sd = js[k] + jd[k];
                xp[k] = x[k] + Dt * (v[k] + Dt * (SECONDORDER * acc_sd + THIRDORDER * Dt * jerk_sd));
                vp[k] = v[k] + Dt * (acc_sd + Dt * SECONDORDER * jerk_sd);
                ad[k] = 0.; jd[k] = 0.;
            }
            star_acc_jerk(P, xp, vp, as, js, nullptr);
            for (int ite = 0; ite < 2; ite++) {
                const double Dt2 = Dt * Dt, Dt_inv = recip_block(Dt), Dt_inv2 = Dt_inv * Dt_inv, Dt_inv3 = Dt_inv2 * Dt_inv;
                for (int k = 0; k < 3; k++) {
                    const double Am_s = as[k] - a0s[k], J0_s = Dt * j0s[k], J1_s = Dt * js[k];
                    const double A2_s = 3. * Am_s - (J1_s + 2. * J0_s), A3_s = -2. * Am_s + (J1_s + J0_s);
                    a2s[k] = 2. * Dt_inv2 * (A2_s + 3. * A3_s);
                    a3s[k] = 6. * Dt_inv3 * A3_s;
                    const double Am_d = ad[k] - a0d[k], J0_d = j0d[k] * Dt, J1_d = jd[k] * Dt;
                    const double A2_d = 3. * Am_d - (J1_d + 2. * J0_d), A3_d = -2. * Am_d + (J1_d + J0_d);
                    a2d[k] = 2. * (A2_d + 3. * A3_d) * Dt_inv2;
                    a3d[k] = 6. * A3_d * Dt_inv3;
                    const double A2_sd = A2_s + A2_d, A3_sd = A3_s + A3_d;
                    x[k] = xp[k] + ONE_SIXTIETH * Dt2 * (5. * A2_sd + 3. * ALPHA_POS * A3_sd);
                    v[k] = vp[k] + ONE_TWELFTH * Dt * (4. * A2_sd + 3. * A3_sd);
                }
                star_acc_jerk(P, x, v, as, js, nullptr);
            }
            time += dt;
            dt = dt_block(P, time, dt, V.acc0[i], ad, jd, a2d, a3d, as, js, a2s, a3s);
        }
        star_acc_jerk(P, x, v, as, js, &phs);
        st3(V.pos, i, x);
        st3(V.vel, i, v);
        st3(V.acc_s, i, as);
        st3(V.jerk_s, i, js);
        st3(V.acc_d, i, ad);
        st3(V.jerk_d, i, jd);
        V.phi_s[i] = phs;
        V.phi_d[i] = 0.;
        V.time[i] = time_c + time;
        V.dt[i] = dt;
    }
    V.n_cluster[i] = 1;
}

// ----------------------------------------------------------------- host

HP make_hp(const gplum_params &p) {
    HP h;
    h.m_sun = p.m_sun; h.eps2 = p.eps2; h.eps2_sun = p.eps2_sun; h.dt_tree = p.dt_tree;
    h.dt_min = p.dt_min; h.eta = p.eta; h.eta_0 = p.eta_0; h.eta_sun = p.eta_sun;
    h.eta_sun0 = p.eta_sun0; h.alpha2 = p.alpha2; h.cut = cutoff_of(p); h.collision = p.collision;
    h.terms_max = TERMS_MAX;
    return h;
}

static HardView hard_view(HardBuffers &hb) {
    HardView H{hb.pos.p, hb.vel.p, hb.acc_s.p, hb.jerk_s.p, hb.acc_d.p, hb.jerk_d.p, hb.xp.p, hb.vp.p,
               hb.a0_s.p, hb.j0_s.p, hb.a0_d.p, hb.j0_d.p, hb.a2_s.p, hb.a3_s.p, hb.a2_d.p, hb.a3_d.p,
               hb.mass.p, hb.rinv.p, hb.time.p, hb.dt.p, hb.acc0.p, hb.r_planet.p, hb.f.p, hb.phi_d.p,
               hb.phi_s.p, hb.id.p, hb.flags.p, hb.twin_of.p, hb.act.p};
    if (hb.crec_cap > 0) {
        H.crec = hb.crec.p;
        H.ccount = hb.ccount.p;
        H.ccap = hb.crec_cap;
    }
    return H;
}


// ------------------------------------------------------------ giant clusters
//
// Clusters above giant_min particles (the reference integrates clusters of 64 and more
// with OpenMP inside the cluster, hard.h:583-658) are spread over the whole GPU: each
// block step is a short sequence of kernels (predict all; collision test over all pairs
// with per-workgroup minima; decision and active flags; the active particles, one wave
// each, as in the multi-wave kernels; the collision aftermath phase by phase; next
// block time and system time), launched in batches by the host, which checks a done
// flag between batches.  Same functions, same summation orders: results identical to the size-class
// kernels.
struct GiantState {
    double t_s, sys_time, ed;
    long long loop, pairs;
    int done, flag_col, ic, jc, n_col;
};

constexpr int GIANT_T = 256;          // threads per workgroup of the grid kernels
constexpr int GIANT_RED = 1024;       // threads of the one-workgroup kernels

__device__ inline void block_min2(double &u, double &v, double *s_u, double *s_v) {
    const int t = threadIdx.x;
    s_u[t] = u;
    s_v[t] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w) {
            if (s_u[t + w] < s_u[t]) s_u[t] = s_u[t + w];
            if (s_v[t + w] < s_v[t]) s_v[t] = s_v[t + w];
        }
        __syncthreads();
    }
    u = s_u[0];
    v = s_v[0];
    __syncthreads();
}

// initial calcGravity of all members with flags cleared (hermite.h:259-270)
__global__ __launch_bounds__(GIANT_T) void k_g_init(HP P, HardView H, int b, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    gravity_full(P, H, b, n, i);
    H.flags[b + i] &= ~(GPLUM_FLAG_DEAD | GPLUM_FLAG_MERGED);
    H.twin_of[b + i] = -1;
}

// loop state and the first block time: t_s = min(time + dt), sys_time = min(time) over
// the live members; done once sys_time reaches the end of the soft step or the cap
__global__ __launch_bounds__(GIANT_RED) void k_g_times(HP P, HardView H, int b, int n, GiantState *G, int first) {
    __shared__ double s_u[GIANT_RED], s_v[GIANT_RED];
    if (!first && G->done) return;
    double tl = 1.e10, tm = 1.e10;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        if (!(H.flags[b + i] & GPLUM_FLAG_DEAD)) {
            tl = dmin(tl, H.time[b + i] + H.dt[b + i]);
            tm = dmin(tm, H.time[b + i]);
        }
    block_min2(tl, tm, s_u, s_v);
    if (threadIdx.x != 0) return;
    if (first) {
        G->loop = 0;
        G->ed = 0.;
        G->n_col = 0;
        G->pairs = (long long)n * (long long)(n - 1);
        G->sys_time = 0.;
    } else {
        G->loop++;
        G->sys_time = tm;
    }
    G->t_s = tl;
    G->flag_col = 0;
    G->done = (G->sys_time < P.dt_tree && G->loop < BLOCK_STEP_CAP) ? 0 : 1;
}

__global__ __launch_bounds__(GIANT_T) void k_g_predict(HardView H, int b, int n, const GiantState *G) {
    if (G->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    predict(H, b + i, G->t_s - H.time[b + i]);
}

// collisionDetermination over the pairs (i, j) of this workgroup: i from blockIdx.x,
// j from the chunk blockIdx.y; per-workgroup minimum in (R, i, j) order, which picks
// the pair the sequential search (ascending i, then j, strict <) finds
__device__ inline bool rij_less(double r, int i, int j, double r2, int i2, int j2) {
    return r < r2 || (r == r2 && (i < i2 || (i == i2 && j < j2)));
}

__global__ __launch_bounds__(GIANT_T) void k_g_collide(HardView H, int b, int n, const GiantState *G,
                                                       double *__restrict__ part_r, int *__restrict__ part_i,
                                                       int *__restrict__ part_j) {
    __shared__ double s_r[GIANT_T];
    __shared__ int s_i[GIANT_T], s_j[GIANT_T];
    if (G->done) return;
    const int t = threadIdx.x;
    const int i = blockIdx.x * blockDim.x + t;
    const int jchunk = (n + gridDim.y - 1) / gridDim.y;
    const int j0 = blockIdx.y * jchunk, j1 = min(n, j0 + jchunk);
    double rmin = 1.;
    int bi = 0x7fffffff, bj = 0x7fffffff;
    if (i < n && !(H.flags[b + i] & GPLUM_FLAG_DEAD)) {
        const double xi = H.xp[3 * (b + i)], yi = H.xp[3 * (b + i) + 1], zi = H.xp[3 * (b + i) + 2];
        const double fri = H.f[b + i] * H.r_planet[b + i];
        for (int j = j0; j < j1; j++) {
            if (i == j || (H.flags[b + j] & GPLUM_FLAG_DEAD)) continue;
            const double dx = xi - H.xp[3 * (b + j)];
            const double dy = yi - H.xp[3 * (b + j) + 1];
            const double dz = zi - H.xp[3 * (b + j) + 2];
            const double d2 = dot3(dx, dy, dz, dx, dy, dz);
            const double r2 = fri + H.f[b + j] * H.r_planet[b + j];
            if (!(d2 < COL_FILTER * r2 * r2)) continue;
            const double Rr = sqrt(d2) / r2;
            if (Rr < rmin) { rmin = Rr; bi = i; bj = j; }
        }
    }
    s_r[t] = rmin; s_i[t] = bi; s_j[t] = bj;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w && rij_less(s_r[t + w], s_i[t + w], s_j[t + w], s_r[t], s_i[t], s_j[t])) {
            s_r[t] = s_r[t + w]; s_i[t] = s_i[t + w]; s_j[t] = s_j[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        const int k = blockIdx.y * gridDim.x + blockIdx.x;
        part_r[k] = s_r[0]; part_i[k] = s_i[0]; part_j[k] = s_j[0];
    }
}

// the pair with the smallest R (< 1: a collision, the lighter body is the impactor; all
// live members are corrected to t_s), then the active flags of the step
__global__ __launch_bounds__(GIANT_RED) void k_g_decide(HP P, HardView H, int b, int n, GiantState *G,
                                                        const double *__restrict__ part_r,
                                                        const int *__restrict__ part_i,
                                                        const int *__restrict__ part_j, int nparts) {
    __shared__ double s_r[GIANT_RED];
    __shared__ int s_i[GIANT_RED], s_j[GIANT_RED];
    __shared__ int s_col;
    if (G->done) return;
    const int t = threadIdx.x;
    double r = 1.;
    int bi = 0x7fffffff, bj = 0x7fffffff;
    if (P.collision)
        for (int k = t; k < nparts; k += blockDim.x)
            if (rij_less(part_r[k], part_i[k], part_j[k], r, bi, bj)) { r = part_r[k]; bi = part_i[k]; bj = part_j[k]; }
    s_r[t] = r; s_i[t] = bi; s_j[t] = bj;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w && rij_less(s_r[t + w], s_i[t + w], s_j[t + w], s_r[t], s_i[t], s_j[t])) {
            s_r[t] = s_r[t + w]; s_i[t] = s_i[t + w]; s_j[t] = s_j[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        s_col = 0;
        if (s_r[0] < 1.) {
            const int gi = s_i[0], gj = s_j[0];
            s_col = 1;
            if (H.mass[b + gi] < H.mass[b + gj]) { G->ic = gi; G->jc = gj; } else { G->ic = gj; G->jc = gi; }
        }
        G->flag_col = s_col;
    }
    __syncthreads();
    const double t_s = G->t_s;
    const int col = s_col;
    for (int i = t; i < n; i += blockDim.x) {
        if (col) H.dt[b + i] = t_s - H.time[b + i];
        H.act[b + i] = !(H.flags[b + i] & GPLUM_FLAG_DEAD) && (col || H.time[b + i] + H.dt[b + i] == t_s);
    }
}

// the active members, one wave each (coop_correct_active over the whole grid)
__global__ __launch_bounds__(GIANT_T) void k_g_correct(HP P, HardView H, int b, int n, GiantState *G) {
    if (G->done) return;
    const JData J{H.xp + 3 * b, H.vp + 3 * b, H.mass + b, H.rinv + b, H.id + b, 3};
    long long pairs = 0;
    const int waves = (int)(blockDim.x >> 6);
    coop_correct_active(P, H, J, b, n, G->t_s, G->flag_col, pairs, (int)blockIdx.x * waves + (int)(threadIdx.x >> 6),
                        (int)gridDim.x * waves, H.twin_of + b);
    if ((threadIdx.x & 63) == 0 && pairs) atomicAdd((unsigned long long *)&G->pairs, (unsigned long long)pairs);
}

// perfect merging and its aftermath (the phases of collision_tail, each over the grid;
// all return at once in the block steps without a collision)
__global__ void k_g_merge(HP P, HardView H, int b, int n, GiantState *G) {
    if (G->done || !G->flag_col || threadIdx.x != 0) return;
    G->ed += perfect_merge(P, H, b, n, G->ic, G->jc);
    G->n_col++;
}
__global__ __launch_bounds__(GIANT_T) void k_g_tail_gravity(HP P, HardView H, int b, int n, GiantState *G) {
    if (G->done || !G->flag_col) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    gravity_full(P, H, b, n, i);
    atomicAdd((unsigned long long *)&G->pairs, (unsigned long long)(n - 1));
}
__global__ __launch_bounds__(GIANT_T) void k_g_tail_merged(HardView H, int b, int n, const GiantState *G) {
    if (G->done || !G->flag_col) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && (H.flags[b + i] & GPLUM_FLAG_MERGED)) merge_acc_jerk(H, b, n, i);
}
__global__ __launch_bounds__(GIANT_T) void k_g_tail_dt(HP P, HardView H, int b, int n, const GiantState *G) {
    if (G->done || !G->flag_col) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int g = b + i;
    if (!(H.flags[g] & GPLUM_FLAG_DEAD)) {
        double ad[3], jd[3], as[3], js[3];
        ld3(H.acc_d, g, ad); ld3(H.jerk_d, g, jd); ld3(H.acc_s, g, as); ld3(H.jerk_s, g, js);
        H.dt[g] = dt_initial(P, H.time[g], H.dt[g], H.acc0[g], ad, jd, as, js);
    }
}
__global__ __launch_bounds__(GIANT_T) void k_g_tail_twins(HardView H, int b, int n, const GiantState *G) {
    if (G->done || !G->flag_col) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && H.twin_of[b + i] >= 0) H.dt[b + i] = H.dt[b + H.twin_of[b + i]];
}

// final calcGravity with potentials (hermite.h:466-470), then the merged targets
__global__ __launch_bounds__(GIANT_T) void k_g_final(HP P, HardView H, int b, int n, GiantState *G) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    gravity_full(P, H, b, n, i);
    atomicAdd((unsigned long long *)&G->pairs, (unsigned long long)(n - 1));
}
__global__ __launch_bounds__(GIANT_T) void k_g_final_merged(HardView H, int b, int n, const GiantState *G, int c,
                                                            double *edisp, int *ncol, long long *nblock,
                                                            long long *npair, double time_end) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && G->n_col > 0 && (H.flags[b + i] & GPLUM_FLAG_MERGED)) merge_acc_jerk(H, b, n, i);
    if (i == 0) {
        edisp[c] = G->ed;
        ncol[c] = G->n_col;
        nblock[c] = (G->sys_time < time_end) ? -1 : G->loop;
        npair[c] = G->pairs;
    }
}

// Giant clusters run concurrently, one per stream of a small per-device pool (streams
// at the default priority, created once and kept for the process).  Each slot captures
// GIANT_BATCH block steps of its cluster into a hipGraph once (10 kernels per step, no
// memset nodes; every kernel returns at once after the done flag is set) and replays it
// per round, then reads its done flag; a finished slot takes the next giant cluster.
// The graph keeps the host from spending its time on kernel launches.
constexpr int GIANT_BATCH = 32;
constexpr int GIANT_STREAMS = 8;

struct GiantSlot {
    int c = -1, b = 0, n = 0, g_n = 0, g_jc = 0, g_w = 0;
    GiantState *G = nullptr;
    double *r = nullptr;
    int *pi = nullptr, *pj = nullptr;
    hipGraphExec_t exec = nullptr;      // GIANT_BATCH block steps of this cluster
};

struct GiantPool {
    hipStream_t st[GIANT_STREAMS];
    hipEvent_t join[GIANT_STREAMS];
    hipEvent_t fork;
    int *done_h;          // pinned host copies of the slots' done flags (asynchronous reads)
};

// the calling thread's device's pool (created on first use)
static int giant_pool(GiantPool *out) {
    static std::mutex mu;
    static GiantPool pool[64] = {};
    int dev = 0;
    GPLUM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(GPLUM_E_NODEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    GiantPool &p = pool[dev];
    for (int k = 0; k < GIANT_STREAMS; k++) {
        if (!p.st[k]) GPLUM_HIP(hipStreamCreateWithFlags(&p.st[k], hipStreamNonBlocking));
        if (!p.join[k]) GPLUM_HIP(hipEventCreateWithFlags(&p.join[k], hipEventDisableTiming));
    }
    if (!p.fork) GPLUM_HIP(hipEventCreateWithFlags(&p.fork, hipEventDisableTiming));
    if (!p.done_h) GPLUM_HIP(hipHostMalloc((void **)&p.done_h, sizeof(int) * GIANT_STREAMS, hipHostMallocDefault));
    *out = p;
    return GPLUM_OK;
}

static void giant_start(const HP &P, HardView &H, GiantSlot &g, hipStream_t st) {
    hipLaunchKernelGGL(k_g_init, dim3(g.g_n), dim3(GIANT_T), 0, st, P, H, g.b, g.n);
    hipLaunchKernelGGL(k_g_times, dim3(1), dim3(GIANT_RED), 0, st, P, H, g.b, g.n, g.G, 1);
}

static void giant_steps(const HP &P, HardView &H, GiantSlot &g, hipStream_t st) {
    for (int s = 0; s < GIANT_BATCH; s++) {
        hipLaunchKernelGGL(k_g_predict, dim3(g.g_n), dim3(GIANT_T), 0, st, H, g.b, g.n, g.G);
        hipLaunchKernelGGL(k_g_collide, dim3(g.g_n, g.g_jc), dim3(GIANT_T), 0, st, H, g.b, g.n, g.G, g.r, g.pi, g.pj);
        hipLaunchKernelGGL(k_g_decide, dim3(1), dim3(GIANT_RED), 0, st, P, H, g.b, g.n, g.G, g.r, g.pi, g.pj,
                           g.g_n * g.g_jc);
        hipLaunchKernelGGL(k_g_correct, dim3(g.g_w), dim3(GIANT_T), 0, st, P, H, g.b, g.n, g.G);
        hipLaunchKernelGGL(k_g_merge, dim3(1), dim3(64), 0, st, P, H, g.b, g.n, g.G);
        hipLaunchKernelGGL(k_g_tail_gravity, dim3(g.g_n), dim3(GIANT_T), 0, st, P, H, g.b, g.n, g.G);
        hipLaunchKernelGGL(k_g_tail_merged, dim3(g.g_n), dim3(GIANT_T), 0, st, H, g.b, g.n, g.G);
        hipLaunchKernelGGL(k_g_tail_dt, dim3(g.g_n), dim3(GIANT_T), 0, st, P, H, g.b, g.n, g.G);
        hipLaunchKernelGGL(k_g_tail_twins, dim3(g.g_n), dim3(GIANT_T), 0, st, H, g.b, g.n, g.G);
        hipLaunchKernelGGL(k_g_times, dim3(1), dim3(GIANT_RED), 0, st, P, H, g.b, g.n, g.G, 0);
    }
}

static void giant_finish(const HP &P, HardView &H, HardBuffers &hb, GiantSlot &g, hipStream_t st) {
    hipLaunchKernelGGL(k_g_final, dim3(g.g_n), dim3(GIANT_T), 0, st, P, H, g.b, g.n, g.G);
    hipLaunchKernelGGL(k_g_final_merged, dim3(g.g_n), dim3(GIANT_T), 0, st, H, g.b, g.n, g.G, g.c, hb.edisp.p,
                       hb.ncol.p, hb.nblock.p, hb.npair.p, P.dt_tree);
}

// Integrates the clusters listed in giants (host offsets off) on the giant path; the
// slots start after event fork (recorded on st once the cluster data were gathered) and
// st waits for all of them.
static int hard_giants(const HP &P, HardView &H, HardBuffers &hb, const std::vector<int> &giants,
                       const std::vector<int32_t> &off, hipStream_t st, const GiantPool &gp) {
    const hipStream_t *se = gp.st;
    int rc = GPLUM_OK;
    const int K = std::min<int>(GIANT_STREAMS, (int)giants.size());
    int max_n = 0;
    for (int c : giants) max_n = std::max(max_n, off[c + 1] - off[c]);
    const int max_parts = ceil_div(max_n, GIANT_T) * std::max(1, std::min(64, ceil_div(max_n, 64)));
    GPLUM_ENSURE(hb.g_state, sizeof(GiantState) * (size_t)K);
    GPLUM_ENSURE(hb.g_r, (size_t)max_parts * K);
    GPLUM_ENSURE(hb.g_i, (size_t)max_parts * K);
    GPLUM_ENSURE(hb.g_j, (size_t)max_parts * K);
    for (int k = 0; k < K; k++) GPLUM_HIP(hipStreamWaitEvent(se[k], gp.fork, 0));
    std::vector<GiantSlot> slot(K);
    size_t next = 0;
    auto release = [&](GiantSlot &g) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
    };
    auto assign = [&](int k) -> int {
        GiantSlot &g = slot[k];
        release(g);
        if (next >= giants.size()) { g.c = -1; return GPLUM_OK; }
        g.c = giants[next++];
        g.b = off[g.c];
        g.n = off[g.c + 1] - g.b;
        g.g_n = ceil_div(g.n, GIANT_T);
        g.g_jc = std::max(1, std::min(64, ceil_div(g.n, 64)));
        g.g_w = std::min(4096, ceil_div(g.n, GIANT_T / 64));
        g.G = reinterpret_cast<GiantState *>(hb.g_state.p) + k;
        g.r = hb.g_r.p + (size_t)k * max_parts;
        g.pi = hb.g_i.p + (size_t)k * max_parts;
        g.pj = hb.g_j.p + (size_t)k * max_parts;
        giant_start(P, H, g, se[k]);
        hipGraph_t graph = nullptr;
        GPLUM_HIP(hipStreamBeginCapture(se[k], hipStreamCaptureModeThreadLocal));
        giant_steps(P, H, g, se[k]);
        const hipError_t ec = hipStreamEndCapture(se[k], &graph);
        if (ec != hipSuccess) return fail(GPLUM_E_HIP, "graph capture of the giant-cluster steps failed");
        const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) return fail(GPLUM_E_HIP, "graph instantiation of the giant-cluster steps failed");
        return GPLUM_OK;
    };
    for (int k = 0; k < K; k++)
        if ((rc = assign(k))) break;
    for (; !rc;) {
        bool any = false;
        for (int k = 0; k < K; k++)
            if (slot[k].c >= 0) {
                GPLUM_HIP(hipGraphLaunch(slot[k].exec, se[k]));
                GPLUM_HIP(hipMemcpyAsync(&gp.done_h[k], &slot[k].G->done, sizeof(int), hipMemcpyDeviceToHost,
                                         se[k]));
                any = true;
            }
        if (!any) break;
        GPLUM_HIP(hipGetLastError());
        for (int k = 0; k < K; k++)
            if (slot[k].c >= 0) {
                GPLUM_HIP(hipStreamSynchronize(se[k]));
                if (gp.done_h[k]) {
                    giant_finish(P, H, hb, slot[k], se[k]);
                    if ((rc = assign(k))) break;
                }
            }
    }
    for (int k = 0; k < K; k++) release(slot[k]);
    if (rc) return rc;
    // st waits for every slot
    for (int k = 0; k < K; k++) {
        GPLUM_HIP(hipEventRecord(gp.join[k], se[k]));
        GPLUM_HIP(hipStreamWaitEvent(st, gp.join[k], 0));
    }
    GPLUM_HIP(hipGetLastError());
    return GPLUM_OK;
}

// zeroes the three cluster-queue counters (a kernel, not a memset, so that the launch
// sequence can be captured into a graph and replayed)
__global__ void k_reset_queue(int *next) {
    if (threadIdx.x < 3) next[threadIdx.x] = 0;
}

int hard_launch(const gplum_params &p, HardBuffers &hb, int ncl, const int32_t *d_off,
                const int32_t *d_order, bool persistent, int n_cu, hipStream_t st, hipStream_t st2,
                hipEvent_t ev_fork, hipEvent_t ev_join, int wg_per_cu, int lds_max, int giant_min, int terms_max) {
    if (ncl <= 0) return GPLUM_OK;
    HP P = make_hp(p);
    P.terms_max = std::max(0, std::min(TERMS_MAX, terms_max));
    HardView H = hard_view(hb);
    GPLUM_ENSURE(hb.next, 3);
    GPLUM_ENSURE(hb.edisp, ncl);
    GPLUM_ENSURE(hb.ncol, ncl);
    GPLUM_ENSURE(hb.nblock, ncl);
    GPLUM_ENSURE(hb.npair, ncl);
    hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, st, hb.next.p);
    const int pers = persistent ? 1 : 0;
    lds_max = std::min(HERMITE_LDS, std::max(HARD_SEQ_MAX, lds_max));
    const int grid = persistent ? std::min(ncl, std::max(1, n_cu) * std::max(1, wg_per_cu)) : ncl;
    // clusters above HARD_SEQ_MAX on a second stream (when given), concurrently with
    // the register-path kernel
    const bool big = hb.max_size > HARD_SEQ_MAX;
    hipStream_t sb = (st2 && big) ? st2 : st;
    if (sb != st) {
        GPLUM_HIP(hipEventRecord(ev_fork, st));
        GPLUM_HIP(hipStreamWaitEvent(sb, ev_fork, 0));
    }
    giant_min = std::max(HARD_SEQ_MAX, giant_min);
    // giant clusters start right after the gather: their stream pool waits on this event
    GiantPool gpool{};
    if (hb.max_size > giant_min) {
        int rc = giant_pool(&gpool);
        if (rc) return rc;
        GPLUM_HIP(hipEventRecord(gpool.fork, st));
    }
    if (hb.max_size > lds_max)
        hipLaunchKernelGGL((k_hermite_multi<false>), dim3(grid), dim3(HERMITE_BIG_BLOCK), 0, sb, P, H, d_off, d_order,
                           ncl, hb.next.p + 2, hb.edisp.p, hb.ncol.p, hb.nblock.p, hb.npair.p, pers, lds_max,
                           giant_min);
    if (big)
        hipLaunchKernelGGL((k_hermite_multi<true>), dim3(grid), dim3(HERMITE_BIG_BLOCK), 0, sb, P, H, d_off, d_order,
                           ncl, hb.next.p + 1, hb.edisp.p, hb.ncol.p, hb.nblock.p, hb.npair.p, pers, lds_max,
                           giant_min);
    hipLaunchKernelGGL((k_hermite<true, true>), dim3(grid), dim3(HERMITE_BLOCK), 0, st, P, H, d_off, d_order, ncl,
                       hb.next.p, hb.edisp.p, hb.ncol.p, hb.nblock.p, hb.npair.p, pers, lds_max, giant_min);
    if (hb.max_size > giant_min) {
        // giant clusters: several at a time over the whole GPU on a stream pool (the
        // size-class kernels above run meanwhile)
        std::vector<int32_t> off(ncl + 1);
        GPLUM_HIP(hipMemcpyAsync(off.data(), d_off, sizeof(int32_t) * (ncl + 1), hipMemcpyDeviceToHost, st));
        GPLUM_HIP(hipStreamSynchronize(st));
        std::vector<int> giants;
        for (int c = 0; c < ncl; c++)
            if (off[c + 1] - off[c] > giant_min) giants.push_back(c);
        // largest first
        std::stable_sort(giants.begin(), giants.end(),
                         [&](int x, int y) { return off[x + 1] - off[x] > off[y + 1] - off[y]; });
        int rc = hard_giants(P, H, hb, giants, off, st, gpool);
        if (rc) return rc;
    }
    if (sb != st) {
        GPLUM_HIP(hipEventRecord(ev_join, sb));
        GPLUM_HIP(hipStreamWaitEvent(st, ev_join, 0));
    }
    GPLUM_HIP(hipGetLastError());
    return GPLUM_OK;
}

// Sun + hard acc/jerk at the current state and the first block step
// (calcGravity + calcDeltatInitial, gravity_hard.h:203-256, particle.h:886-914) of
// every particle of clusters [d_off[c], d_off[c+1]): one 64-thread block per cluster.
__global__ __launch_bounds__(HERMITE_BLOCK) void k_hard_prepare(HP P, HardView H, const int32_t *__restrict__ d_off,
                                                                int ncl) {
    const int c = blockIdx.x;
    if (c >= ncl) return;
    const int b = d_off[c], n = d_off[c + 1] - b;
    for (int i = (int)threadIdx.x; i < n; i += HERMITE_BLOCK) {
        gravity_full(P, H, b, n, i);
        double ad[3], jd[3], as[3], js[3];
        ld3(H.acc_d, b + i, ad); ld3(H.jerk_d, b + i, jd);
        ld3(H.acc_s, b + i, as); ld3(H.jerk_s, b + i, js);
        H.dt[b + i] = dt_initial(P, H.time[b + i], H.dt[b + i], H.acc0[b + i], ad, jd, as, js);
    }
}

int hard_prepare(const gplum_params &p, HardBuffers &hb, int ncl, const int32_t *d_off, hipStream_t st) {
    if (ncl <= 0) return GPLUM_OK;
    HP P = make_hp(p);
    HardView H{hb.pos.p, hb.vel.p, hb.acc_s.p, hb.jerk_s.p, hb.acc_d.p, hb.jerk_d.p, hb.xp.p, hb.vp.p,
               hb.a0_s.p, hb.j0_s.p, hb.a0_d.p, hb.j0_d.p, hb.a2_s.p, hb.a3_s.p, hb.a2_d.p, hb.a3_d.p,
               hb.mass.p, hb.rinv.p, hb.time.p, hb.dt.p, hb.acc0.p, hb.r_planet.p, hb.f.p, hb.phi_d.p,
               hb.phi_s.p, hb.id.p, hb.flags.p, hb.twin_of.p, hb.act.p};
    hipLaunchKernelGGL(k_hard_prepare, dim3(ncl), dim3(HERMITE_BLOCK), 0, st, P, H, d_off, ncl);
    GPLUM_HIP(hipGetLastError());
    return GPLUM_OK;
}

int hard_launch_single(const gplum_params &p, HardBuffers &hb, int c, int n_c, const int32_t *d_off,
                       hipStream_t st) {
    // one launch integrating cluster c (n_c particles) only: the per-cluster-launch comparator
    HP P = make_hp(p);
    HardView H = hard_view(hb);
    if (n_c <= HARD_SEQ_MAX)
        hipLaunchKernelGGL((k_hermite<true, true>), dim3(1), dim3(HERMITE_BLOCK), 0, st, P, H, d_off + c, nullptr, 1,
                           hb.next.p, hb.edisp.p + c, hb.ncol.p + c, hb.nblock.p + c, hb.npair.p + c, 0, HERMITE_LDS, 0x7fffffff);
    else if (n_c <= HERMITE_LDS)
        hipLaunchKernelGGL((k_hermite_multi<true>), dim3(1), dim3(HERMITE_BIG_BLOCK), 0, st, P, H, d_off + c, nullptr, 1,
                           hb.next.p, hb.edisp.p + c, hb.ncol.p + c, hb.nblock.p + c, hb.npair.p + c, 0, HERMITE_LDS, 0x7fffffff);
    else
        hipLaunchKernelGGL((k_hermite_multi<false>), dim3(1), dim3(HERMITE_BIG_BLOCK), 0, st, P, H, d_off + c, nullptr,
                           1, hb.next.p, hb.edisp.p + c, hb.ncol.p + c, hb.nblock.p + c, hb.npair.p + c, 0, HERMITE_LDS, 0x7fffffff);
    GPLUM_HIP(hipGetLastError());
    return GPLUM_OK;
}

int isolated_launch(const gplum_params &p, const IsoArgs &a, int n, double t0, double t1, hipStream_t st) {
    if (n <= 0) return GPLUM_OK;
    HP P = make_hp(p);
    IsoView V{a.pos, a.vel, a.acc_d, a.jerk_d, a.acc_s, a.jerk_s, a.phi_d, a.phi_s, a.time, a.dt, a.acc0,
              a.number, a.n_cluster};
    hipLaunchKernelGGL(k_isolated, dim3(ceil_div(n, 128)), dim3(128), 0, st, P, V, n, t0, t1);
    GPLUM_HIP(hipGetLastError());
    return GPLUM_OK;
}


// ------------------------------------------------------- time-step self-check

__device__ inline uint64_t mix64(uint64_t z) {     // splitmix64 finaliser
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
// sign * [0.5, 1) * 2^e with e uniform in [e0, e1]
__device__ inline double rnd_mag(uint64_t &st, int e0, int e1) {
    st += 0x9E3779B97F4A7C15ULL;
    const uint64_t r = mix64(st);
    const double m = 0.5 + 0.5 * (double)(r >> 11) * 0x1p-53;
    const int e = e0 + (int)(mix64(r ^ 0x5851F42D4C957F2DULL) % (uint64_t)(e1 - e0 + 1));
    return ((r >> 10) & 1) ? -ldexp(m, e) : ldexp(m, e);
}

// dt_block with the fast criterion against dt_block with calc_dt4 on random states:
// accelerations and their derivatives over many decades, steps dt_tree 2^-k and times on
// the step's grid.  One sample in four widens every magnitude by 2^+-540, so that the
// products of squares in the fast criterion underflow below DBL_MIN or overflow past
// DBL_MAX.  Counts results that differ in any bit and fast values sent to the exact
// evaluation.
__global__ void k_dt_selfcheck(HP P, long long n, uint64_t seed, unsigned long long *out) {
    unsigned long long mism = 0, fb = 0;
    for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < n; s += (long long)gridDim.x * blockDim.x) {
        uint64_t st = seed ^ mix64((uint64_t)s * 0x2545F4914F6CDD1DULL);
        double v[8][3];
        const int e_lo[8] = {-20, -24, -28, -32, -12, -16, -20, -24};
        const int e_hi[8] = {8, 14, 20, 26, 4, 8, 12, 16};
        const int wide = ((s & 3) == 3) ? 540 : 0;
        for (int q = 0; q < 8; q++)
            for (int k = 0; k < 3; k++) v[q][k] = rnd_mag(st, e_lo[q] - wide, e_hi[q] + wide);
        const double acc0 = fabs(rnd_mag(st, -12, 4));
        st += 0x9E3779B97F4A7C15ULL;
        const uint64_t r = mix64(st);
        const int kk = 1 + (int)(r % 30);                       // dt = dt_tree 2^-kk
        const double dt = ldexp(P.dt_tree, -kk);
        const double time = dt * (double)((r >> 8) % (1ULL << kk));   // on the grid of dt
        int fell = 0;
        const double a = dt_block<true>(P, time + dt, dt, acc0, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], &fell);
        const double b = dt_block<false>(P, time + dt, dt, acc0, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]);
        mism += (__double_as_longlong(a) != __double_as_longlong(b)) ? 1 : 0;
        fb += fell;
    }
    atomicAdd(&out[0], mism);
    atomicAdd(&out[1], fb);
}

int hermite_dt_selfcheck(const gplum_params &p, long long n, uint64_t seed, long long *mismatches,
                         long long *fallbacks) {
    if (n < 0) return fail(GPLUM_E_ARG, "negative sample count");
    DArray<unsigned long long> cnt;
    GPLUM_ENSURE(cnt, 2);
    GPLUM_HIP(hipMemset(cnt.p, 0, 2 * sizeof(unsigned long long)));
    hipLaunchKernelGGL(k_dt_selfcheck, dim3(1024), dim3(256), 0, 0, make_hp(p), n, seed, cnt.p);
    GPLUM_HIP(hipGetLastError());
    unsigned long long h[2] = {0, 0};
    GPLUM_HIP(hipMemcpy(h, cnt.p, sizeof(h), hipMemcpyDeviceToHost));
    if (mismatches) *mismatches = (long long)h[0];
    if (fallbacks) *fallbacks = (long long)h[1];
    return GPLUM_OK;
}

}  // namespace gplum
