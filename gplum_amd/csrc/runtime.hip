// runtime.hip -- error state, parameter defaults and device queries of the C ABI.
#include <cmath>
#include <cstring>
#include <string>

#include "common.hpp"

namespace gplum {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

}  // namespace gplum

extern "C" {

const char *gplum_last_error(void) { return gplum::g_last_error.c_str(); }

// Reference defaults: static member initialisers of src/particle.h:917-961 and the
// tree parameters of src/read.h (theta 0.5, n_leaf_limit 8, n_group_limit 256).
void gplum_params_default(gplum_params *p) {
    std::memset(p, 0, sizeof(*p));
    p->m_sun = 1.;
    p->dens = 5.049667e6;
    p->R_cut0 = 2.;
    p->R_cut1 = 8.;
    p->R_search0 = 1.;
    p->R_search1 = 4.;
    p->R_search2 = 1.;
    p->R_search3 = 4.;
    p->increase_factor = 1.;
    p->dt_tree = 1.0 / 32.0;
    p->dt_min = 1.0 / 8192.0;
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
    gplum_params_set_gamma(p, 0.5);
}

// FPGrav::setGamma, src/particle.h:619-633.
void gplum_params_set_gamma(gplum_params *p, double g) {
    p->gamma = g;
    p->g_1_inv = 1. / (g - 1.);
    const double g2 = g * g;
    const double g_1_inv3 = p->g_1_inv * p->g_1_inv * p->g_1_inv;
    p->g_1_inv7 = g_1_inv3 * g_1_inv3 * p->g_1_inv;
    p->w_y = 7. / 3. * ((((((g - 9.) * g + 45.) * g - 60. * std::log(g)) * g - 45.) * g + 9.) * g - 1.) *
             p->g_1_inv7;
    p->f1 = (-10. / 3. + 14. * (g + 1.) - 21. * ((g + 3.) * g + 1.) +
             35. / 3. * (((g + 9.) * g + 9.) * g + 1.) - 70. * ((g + 3.) * g + 1.) * g +
             210. * (g + 1.) * g2 + (((g - 7.) * g + 21.) * g - 35.) * g2 * g2) *
            p->g_1_inv7;
}

int gplum_device_info(char *name, int name_len, int *n_cu, double *clock_ghz) {
    int dev = 0, count = 0;
    GPLUM_HIP(hipGetDeviceCount(&count));
    if (count <= 0) return gplum::fail(GPLUM_E_NODEVICE, "no HIP device");
    GPLUM_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    GPLUM_HIP(hipGetDeviceProperties(&prop, dev));
    if (name && name_len > 0) {
        std::snprintf(name, (size_t)name_len, "%s", prop.gcnArchName);
    }
    if (n_cu) *n_cu = prop.multiProcessorCount;
    if (clock_ghz) *clock_ghz = prop.clockRate * 1e-6;
    return GPLUM_OK;
}

}  // extern "C"
