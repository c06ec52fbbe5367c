// fdps_functor_demo.cpp -- drives the reference's functor interface (gplum_fdps.hpp)
// the way FDPS would, with user types laid out like src/particle.h:16-170, and checks
// the results against a direct FP64 evaluation of the same interactions.
//
//   g++ -std=c++14 -O2 -I include fdps_functor_demo.cpp -L gplum_amd/lib -lgplum_gpu
// Exit status 0 and "fdps functors: OK" on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "gplum_fdps.hpp"

// user-side types with the reference layouts (individual cut-off build)
struct NeighborInfo { int number, rank, id_max, id_min; };
struct ForceGrav {
    float acc[3];
    float phi;
    NeighborInfo neighbor;
    void clear() {
        acc[0] = acc[1] = acc[2] = 0.f; phi = 0.f;
        neighbor.number = neighbor.rank = 0; neighbor.id_max = -1; neighbor.id_min = 2147483647;
    }
};
struct EPIGrav { int id_local, myrank; double pos[3]; double r_out, r_search; };
struct EPJGrav : EPIGrav { long long id; double mass; double vel[3]; double acc_d[3]; };
struct SPJQuad { double mass; double pos[3]; double quad[6]; };
struct FPGrav { static double eps2; };
double FPGrav::eps2 = 0.;

int main() {
    std::mt19937_64 rng(12345);
    std::normal_distribution<double> nd(0., 1e-3);
    const int ni = 150, nj = 700, ns = 300;
    std::vector<EPIGrav> epi(ni);
    std::vector<EPJGrav> epj(nj);
    std::vector<SPJQuad> spj(ns);
    for (int i = 0; i < nj; i++) {
        EPJGrav &q = epj[i];
        q.id_local = i; q.myrank = 0; q.id = i; q.mass = 1e-10;
        for (int k = 0; k < 3; k++) { q.pos[k] = (k == 0 ? 1.0 : 0.0) + nd(rng); q.vel[k] = 0.; q.acc_d[k] = 0.; }
        q.r_out = 2e-4; q.r_search = 3e-4;
        if (i < ni) epi[i] = q;
    }
    for (int s = 0; s < ns; s++) {
        spj[s].mass = 1e-8;
        for (int k = 0; k < 3; k++) spj[s].pos[k] = (k == 0 ? 1.0 : 0.0) + 50. * nd(rng);
        for (int k = 0; k < 6; k++) spj[s].quad[k] = 0.;
    }
    // per-group calls, as FDPS calls the user functors
    std::vector<ForceGrav> f(ni);
    for (auto &x : f) x.clear();
    gplum_fdps::calcForceEPEPWithSearch<FPGrav>()(epi.data(), ni, epj.data(), nj, f.data());
    gplum_fdps::calcForceEPSP<FPGrav>()(epi.data(), ni, spj.data(), ns, f.data());
    // direct FP64 evaluation of the same sums
    double err2 = 0.;
    int bad_ngb = 0;
    for (int i = 0; i < ni; i++) {
        double a[3] = {0., 0., 0.};
        int ngb = 0;
        for (int j = 0; j < nj; j++) {
            double d[3], r2 = 0.;
            for (int k = 0; k < 3; k++) { d[k] = epj[j].pos[k] - epi[i].pos[k]; r2 += d[k] * d[k]; }
            const double ro = std::max(epi[i].r_out, epj[j].r_out), rs = std::max(epi[i].r_search, epj[j].r_search);
            if (r2 < 1.0201 * rs * rs && j != i) ngb++;
            const double rc2 = std::max(r2, ro * ro);
            const double w = epj[j].mass / (rc2 * std::sqrt(rc2));
            for (int k = 0; k < 3; k++) a[k] += w * d[k];
        }
        for (int s = 0; s < ns; s++) {
            double d[3], r2 = 0.;
            for (int k = 0; k < 3; k++) { d[k] = spj[s].pos[k] - epi[i].pos[k]; r2 += d[k] * d[k]; }
            const double w = spj[s].mass / (r2 * std::sqrt(r2));
            for (int k = 0; k < 3; k++) a[k] += w * d[k];
        }
        double num = 0., den = 0.;
        for (int k = 0; k < 3; k++) { num += (f[i].acc[k] - a[k]) * (f[i].acc[k] - a[k]); den += a[k] * a[k]; }
        err2 += num / den;
        if (ngb != f[i].neighbor.number) bad_ngb++;
    }
    const double rms = std::sqrt(err2 / ni);
    // multi-walk: three walks (the whole group, and two halves), compared with the per-call result
    const EPIGrav *pe[3] = {epi.data(), epi.data(), epi.data() + ni / 2};
    const int ne[3] = {ni, ni / 2, ni - ni / 2};
    const EPJGrav *pj[3] = {epj.data(), epj.data(), epj.data()};
    const int nje[3] = {nj, nj, nj};
    const SPJQuad *ps[3] = {spj.data(), spj.data(), spj.data()};
    const int nse[3] = {ns, ns, ns};
    gplum_multiwalk_set_eps2(0.f);
    gplum_fdps::DispatchKernel(0, 3, pe, ne, pj, nje, ps, nse);
    std::vector<ForceGrav> w0(ni), w1(ni / 2), w2(ni - ni / 2);
    ForceGrav *pf[3] = {w0.data(), w1.data(), w2.data()};
    gplum_fdps::RetrieveKernel(0, 3, ne, pf);
    int mw_bad = 0;
    for (int i = 0; i < ni; i++)
        for (int k = 0; k < 3; k++)
            if (std::fabs(w0[i].acc[k] - f[i].acc[k]) > 1e-6 * std::fabs(f[i].acc[k]) + 1e-30) mw_bad++;
    for (int i = 0; i < ni; i++)
        if (w0[i].neighbor.number != f[i].neighbor.number) mw_bad++;
    std::printf("fdps functors: rms rel acc err vs FP64 direct %.3e, neighbour-count mismatches %d, "
                "multiwalk mismatches %d\n", rms, bad_ngb, mw_bad);
    if (rms > 1e-6 || bad_ngb != 0 || mw_bad != 0) {
        std::printf("fdps functors: FAIL\n");
        return 1;
    }
    std::printf("fdps functors: OK\n");
    return 0;
}
