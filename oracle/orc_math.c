/* orc_math.c -- transcendental functions built from IEEE-exact operations only
 * (frexp, ldexp, rint, +, -, *, /), so that the oracle and the HIP kernels
 * (gplum_amd/csrc/common.hpp, same operation sequence) produce identical bits.  Test
 * infrastructure, like the rest of oracle/.
 *
 * They stand in for the library calls of the reference where the last bit of the
 * result decides something later: pow(x, 1./3.) for the merged radius
 * (collisionA.h:251) and the Hill radius (particle.h:694-699), exp() in the velocity
 * dispersion smoothing (particle.h:1436-1449).  Each is within 1 ulp of the true
 * value (tests/test_oracle_hermite.py, tests/test_oracle.py). */
#include <math.h>
#include "oracle.h"

/* x^(1/3), x > 0: quadratic first guess on [0.5, 4) and five Newton steps. */
double orc_cbrt_exact_ops(double x)
{
    int e = 0;
    const double f = frexp(x, &e);
    const int k = (e >= 0) ? e / 3 : -((2 - e) / 3);
    const double u = ldexp(f, e - 3 * k);
    double y = (-0.0368 * u + 0.3769) * u + 0.6515;
    for (int it = 0; it < 5; it++) y = y - (y * y * y - u) / (3.0 * y * y);
    return ldexp(y, k);
}

/* e^x: k = rint(x / ln 2), r = x - k ln 2 (Cody-Waite, exact high part), Taylor
 * series of e^r to degree 13 in Horner form (coefficients 1/m! rounded to double, one
 * multiply and one add per degree), times 2^k. */
static const double ORC_EXP_C[14] = {1.0, 1.0, 0.5, 0.16666666666666666, 0.041666666666666664, 0.008333333333333333, 0.001388888888888889, 0.0001984126984126984, 2.48015873015873e-05, 2.7557319223985893e-06, 2.755731922398589e-07, 2.505210838544172e-08, 2.08767569878681e-09, 1.6059043836821613e-10};
double orc_exp_exact_ops(double x)
{
    if (x < -745.2) return 0.;
    if (x > 709.78) return HUGE_VAL;
    const double LN2_HI = 6.93147180369123816490e-01, LN2_LO = 1.90821492927058770002e-10;
    const double INV_LN2 = 1.44269504088896338700e+00;
    const double k = rint(x * INV_LN2);
    const double r = (x - k * LN2_HI) - k * LN2_LO;
    double s = ORC_EXP_C[13];
    for (int m = 12; m >= 0; m--) s = ORC_EXP_C[m] + r * s;
    return ldexp(s, (int)k);
}
