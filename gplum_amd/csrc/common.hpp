// common.hpp -- shared host/device helpers of libgplum_gpu (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/gplum_gpu.h"
#include "../../include/gplum_system.h"
#include "../../include/gplum_forecast.h"

static_assert(sizeof(gplum_neighbor_info) == 16, "NeighborInfo layout");
static_assert(sizeof(gplum_force) == 32, "ForceGrav layout");
static_assert(sizeof(gplum_epi) == 48, "EPIGrav layout");
static_assert(sizeof(gplum_epj) == 112, "EPJGrav layout");
static_assert(sizeof(gplum_spj_mono) == 32, "SPJMonopole layout");
static_assert(sizeof(gplum_spj_quad) == 80, "SPJQuadrupole layout");

namespace gplum {

// ---------------------------------------------------------------- error state
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define GPLUM_HIP(call)                                                              \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess)                                                        \
            return ::gplum::fail(GPLUM_E_HIP, std::string(#call " -> ") +            \
                                                  hipGetErrorString(e_));            \
    } while (0)

// Bounds on data-dependent loops in device code.  The reference's step-halving
// loops (particle.h:902-912, 1304-1311) terminate only for power-of-two times; the
// caps keep a non-dyadic or NaN input from spinning a wave forever (the result is
// then dt_min, and the block-step cap raises an error on the host).
constexpr int HALVE_CAP = 160;
constexpr long long BLOCK_STEP_CAP = 1ll << 22;

// Clear values of NeighborInfo: S32_MAX-1 with S32_MAX = 2^31 (main_p3t.cpp:32).
constexpr int32_t ID_MIN_CLEAR = 2147483647;
constexpr int32_t ID_MAX_CLEAR = -1;

// ------------------------------------------------------------ device helpers
// std::min / std::max semantics: min(a,b) = (b<a)?b:a, max(a,b) = (a<b)?b:a.
__host__ __device__ inline double smin(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ inline double smax(double a, double b) { return (a < b) ? b : a; }

// fmod(x, d) for x >= 0.  When d is a power of two (every block step is: dt_tree is
// one and the steps are its halvings), x / d is an exact scaling, so
// x - floor(x / d) * d is exact and equals fmod bit for bit at a fraction of its cost;
// other divisors take fmod itself.  The scaling is a multiplication by 2^(1-e)
// (d = 2^(e-1)): x * 2^(1-e) and x / d round the same real number, so they are equal,
// and the multiplication skips the division sequence.
__host__ __device__ inline double fmod_block(double x, double d) {
    int e = 0;
    if (x >= 0. && d > 0. && frexp(d, &e) == 0.5) {
        const double q = (e > -1000) ? x * ldexp(1.0, 1 - e) : x / d;
        return x - floor(q) * d;
    }
    return fmod(x, d);
}

// 1 / d, by a scaling when d is a power of two (the same correctly rounded value as the
// division, as in fmod_block), else by the division.
__host__ __device__ inline double recip_block(double d) {
    int e = 0;
    if (d > 0. && frexp(d, &e) == 0.5 && e > -1000) return ldexp(1.0, 1 - e);
    return 1. / d;
}

// x^(1/3) for finite x > 0 from IEEE-exact operations only (frexp, ldexp, +, -, *, /):
// a quadratic first guess on [0.5, 4) and five Newton steps, within 1 ulp of the true
// root.  The reference calls pow(x, 1./3.) (collisionA.h:251); library pow differs
// by an ulp between the host libm and the device math library, and the merged radius
// decides later collisions, so the oracle and the kernels share this function.
__host__ __device__ inline double cbrt_exact_ops(double x) {
    int e = 0;
    const double f = frexp(x, &e);
    const int k = (e >= 0) ? e / 3 : -((2 - e) / 3);
    const double u = ldexp(f, e - 3 * k);
    double y = (-0.0368 * u + 0.3769) * u + 0.6515;
    for (int it = 0; it < 5; it++) y = y - (y * y * y - u) / (3.0 * y * y);
    return ldexp(y, k);
}

// exp(x) from IEEE-exact operations only (rint, ldexp, +, -, *): k = rint(x / ln 2),
// r = x - k ln2 in two parts (Cody-Waite; the high part has 21 trailing zero bits,
// so k * LN2_HI is exact), e^r by its Taylor series to degree 13 (|r| <= 0.347:
// truncation below 1e-17), scaled by 2^k; within a few ulp of the true value.
// Shared with the oracle so that v_disp (a sum of 32 exponentials per particle,
// particle.h:1436-1449) and the neighbour radii built on it agree bit for bit.
__host__ __device__ inline double exp_exact_ops(double x) {
    if (x < -745.2) return 0.;
    if (x > 709.78) return __builtin_huge_val();
    const double LN2_HI = 6.93147180369123816490e-01, LN2_LO = 1.90821492927058770002e-10;
    const double INV_LN2 = 1.44269504088896338700e+00;
    const double k = rint(x * INV_LN2);
    const double r = (x - k * LN2_HI) - k * LN2_LO;
    // Taylor series of e^r to degree 13, Horner form with the coefficients 1/m! rounded
    // to double (one multiply and one add per degree, no division)
    constexpr double C[14] = {1.0, 1.0, 0.5, 0.16666666666666666, 0.041666666666666664, 0.008333333333333333, 0.001388888888888889, 0.0001984126984126984, 2.48015873015873e-05, 2.7557319223985893e-06, 2.755731922398589e-07, 2.505210838544172e-08, 2.08767569878681e-09, 1.6059043836821613e-10};
    double s = C[13];
#pragma unroll
    for (int m = 12; m >= 0; m--) s = C[m] + r * s;
    return ldexp(s, (int)k);
}

// Exact-order FP64 helpers (the library is compiled with -ffp-contract=off, so
// these are plain IEEE mul/add in the written order).
__host__ __device__ inline double dot3(double ax, double ay, double az,
                                       double bx, double by, double bz) {
    return ax * bx + ay * by + az * bz;
}

// Changeover functions, src/cutfunc.h:4-66 and FPGrav::setGamma constants.
struct Cutoff {
    double gamma, g_1_inv, g_1_inv7, w_y, f1;
};

__host__ __device__ inline double cutoff_f(const Cutoff &c, double y) {
    const double g = c.gamma;
    const double g2 = g * g;
    return (((((((-10. / 3. * y + 14. * (g + 1.)) * y - 21. * ((g + 3.) * g + 1.)) * y
                + 35. / 3. * (((g + 9.) * g + 9.) * g + 1.)) * y
               - 70. * ((g + 3.) * g + 1.) * g) * y
              + 210. * (g + 1.) * g2) * y - 140. * g2 * g * log(y)) * y
            + (((g - 7.) * g + 21.) * g - 35.) * g2 * g2) * c.g_1_inv7;
}
__host__ __device__ inline double cutoff_W(const Cutoff &c, double rij, double r_out_inv) {
    const double y = rij * r_out_inv;
    if (1.0 <= y) return 1.0;
    if (y <= c.gamma) return y * c.w_y;
    return cutoff_f(c, y) + y * (1. - c.f1);
}
__host__ __device__ inline double cutoff_Kx(double x) {
    if (x < 0.) return 0.;
    if (x >= 1.) return 1.;
    const double x2 = x * x;
    return (((-20. * x + 70.) * x - 84.) * x + 35.) * x2 * x2;
}
__host__ __device__ inline double cutoff_K(const Cutoff &c, double rij, double r_out_inv) {
    return cutoff_Kx((c.gamma - rij * r_out_inv) * c.g_1_inv);
}
__host__ __device__ inline double cutoff_dKdt(const Cutoff &c, double rij, double r_out_inv,
                                              double alpha) {
    const double x = (c.gamma - rij * r_out_inv) * c.g_1_inv;
    const double x_1 = x - 1.;
    const double dKdr = (x < 0. || x >= 1.) ? 0.
                        : (140. * x * x * x * x_1 * x_1 * x_1 * r_out_inv * c.g_1_inv);
    return alpha * rij * dKdr;
}

inline Cutoff cutoff_of(const gplum_params &p) {
    return Cutoff{p.gamma, p.g_1_inv, p.g_1_inv7, p.w_y, p.f1};
}

// Wave-level helpers (wave64).
__device__ inline int lane_id() { return __lane_id(); }

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace gplum
