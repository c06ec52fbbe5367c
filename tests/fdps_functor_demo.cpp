// fdps_functor_demo.cpp -- drives the reference's functor interface (gplum_fdps.hpp)
// the way FDPS would, with user types laid out like src/particle.h:16-170, and checks
// the results against a direct FP64 evaluation of the same interactions.  The
// MultiWalk dispatch is checked against the per-group functors (quadrupole and
// monopole super-particles), and the MultiWalk Index dispatch against the list
// dispatch of the same walks (identical bits: the kernel reads the same sequence).
//
//   g++ -std=c++14 -O2 -I include fdps_functor_demo.cpp -L gplum_amd/lib -lgplum_gpu
// Exit status 0 and "fdps functors: OK" on success.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
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
struct SPJMono { double mass; double pos[3]; };

// the walks of the MultiWalk checks: walk 0 holds every i-particle, walks 1 and 2 the
// two halves; walk w's EP / SP lists are the entries k of epj / spj with sel(w, k)
static bool sel_j(int w, int k, int nj) { return w == 0 || (w == 1 ? k < nj / 2 : (k % 2) == 0); }
static bool sel_s(int w, int k, int ns) { return w == 0 || (w == 1 ? k < ns / 3 : (k % 2) == 1); }

template <class SP>
static int multiwalk_index_mismatches(const std::vector<EPIGrav> &epi, const std::vector<EPJGrav> &epj,
                                      const std::vector<SP> &spj, int tag_list, int tag_index, std::mt19937_64 &rng) {
    const int ni = (int)epi.size(), nj = (int)epj.size(), ns = (int)spj.size();
    // whole arrays: epj and spj permuted; pj[k] / ps[k] = where entry k went
    std::vector<int> pj(nj), ps(ns);
    std::iota(pj.begin(), pj.end(), 0);
    std::iota(ps.begin(), ps.end(), 0);
    std::shuffle(pj.begin(), pj.end(), rng);
    std::shuffle(ps.begin(), ps.end(), rng);
    std::vector<EPJGrav> whole_j(nj);
    std::vector<SP> whole_s(ns);
    for (int k = 0; k < nj; k++) whole_j[pj[k]] = epj[k];
    for (int k = 0; k < ns; k++) whole_s[ps[k]] = spj[k];
    std::vector<EPJGrav> lj[3];
    std::vector<SP> ls[3];
    std::vector<int> ij[3], is[3];
    for (int w = 0; w < 3; w++) {
        for (int k = 0; k < nj; k++)
            if (sel_j(w, k, nj)) { lj[w].push_back(epj[k]); ij[w].push_back(pj[k]); }
        for (int k = 0; k < ns; k++)
            if (sel_s(w, k, ns)) { ls[w].push_back(spj[k]); is[w].push_back(ps[k]); }
    }
    const EPIGrav *pe[3] = {epi.data(), epi.data(), epi.data() + ni / 2};
    const int ne[3] = {ni, ni / 2, ni - ni / 2};
    const EPJGrav *pl[3] = {lj[0].data(), lj[1].data(), lj[2].data()};
    const SP *psl[3] = {ls[0].data(), ls[1].data(), ls[2].data()};
    const int *pij[3] = {ij[0].data(), ij[1].data(), ij[2].data()};
    const int *pis[3] = {is[0].data(), is[1].data(), is[2].data()};
    int nje[3], nse[3];
    for (int w = 0; w < 3; w++) { nje[w] = (int)lj[w].size(); nse[w] = (int)ls[w].size(); }
    gplum_fdps::DispatchKernel(tag_list, 3, pe, ne, pl, nje, psl, nse);
    // the whole arrays are sent with the first index dispatch; a second one (walk 2
    // alone, another tag) reuses them
    gplum_fdps::DispatchKernelIndex(tag_index, 3, pe, ne, pij, nje, pis, nse, whole_j.data(), nj, whole_s.data(), ns,
                                    true);
    gplum_fdps::DispatchKernelIndex(tag_index + 1, 1, pe + 2, ne + 2, pij + 2, nje + 2, pis + 2, nse + 2,
                                    (const EPJGrav *)nullptr, 0, (const SP *)nullptr, 0, false);
    std::vector<ForceGrav> a[3], b[3], c2(ne[2]);
    ForceGrav *fa[3], *fb[3], *fc[1] = {c2.data()};
    for (int w = 0; w < 3; w++) {
        a[w].resize(ne[w]); b[w].resize(ne[w]);
        fa[w] = a[w].data(); fb[w] = b[w].data();
    }
    gplum_fdps::RetrieveKernel(tag_list, 3, ne, fa);
    gplum_fdps::RetrieveKernel(tag_index, 3, ne, fb);
    gplum_fdps::RetrieveKernel(tag_index + 1, 1, ne + 2, fc);
    int bad = 0;
    for (int w = 0; w < 3; w++)
        bad += std::memcmp(a[w].data(), b[w].data(), sizeof(ForceGrav) * ne[w]) != 0;
    bad += std::memcmp(a[2].data(), c2.data(), sizeof(ForceGrav) * ne[2]) != 0;
    return bad;
}
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
    // monopole super-particles: per-group functors and the MultiWalk dispatch
    std::vector<SPJMono> spm(ns);
    for (int s = 0; s < ns; s++) {
        spm[s].mass = spj[s].mass;
        for (int k = 0; k < 3; k++) spm[s].pos[k] = spj[s].pos[k];
    }
    std::vector<ForceGrav> fm(ni), wm(ni);
    for (auto &x : fm) x.clear();
    gplum_fdps::calcForceEPEPWithSearch<FPGrav>()(epi.data(), ni, epj.data(), nj, fm.data());
    gplum_fdps::calcForceEPSP<FPGrav>()(epi.data(), ni, spm.data(), ns, fm.data());
    const SPJMono *pm[1] = {spm.data()};
    const EPIGrav *pe0[1] = {epi.data()};
    const EPJGrav *pj0[1] = {epj.data()};
    gplum_fdps::DispatchKernel(1, 1, pe0, &ni, pj0, &nj, pm, &ns);
    ForceGrav *pwm[1] = {wm.data()};
    gplum_fdps::RetrieveKernel(1, 1, &ni, pwm);
    int mono_bad = 0;
    for (int i = 0; i < ni; i++) {
        for (int k = 0; k < 3; k++)
            if (std::fabs(wm[i].acc[k] - fm[i].acc[k]) > 1e-6 * std::fabs(fm[i].acc[k]) + 1e-30) mono_bad++;
        if (wm[i].neighbor.number != fm[i].neighbor.number) mono_bad++;
    }
    // MultiWalk Index against the list dispatch of the same walks, both SPJ types
    const int idx_bad = multiwalk_index_mismatches(epi, epj, spj, 2, 3, rng) +
                        multiwalk_index_mismatches(epi, epj, spm, 5, 6, rng);
    // argument errors of the Index dispatch: an index outside the whole array, and a
    // dispatch without send_flag whose SPJ type differs from the one last sent
    int err_bad = 0;
    {
        const gplum_epi *e1[1] = {reinterpret_cast<const gplum_epi *>(epi.data())};
        const int bad_idx[1] = {nj};
        const int32_t *ij1[1] = {bad_idx}, *is1[1] = {bad_idx};
        const int one = 1, zero = 0;
        if (gplum_multiwalk_dispatch_index_mono(8, 1, e1, &ni, ij1, &one, is1, &zero, nullptr, 0, nullptr, 0, 0) !=
            GPLUM_E_ARG)
            err_bad++;
        const int ok_idx[1] = {0};
        const int32_t *ij2[1] = {ok_idx};
        if (gplum_multiwalk_dispatch_index(8, 1, e1, &ni, ij2, &one, is1, &zero, nullptr, 0, nullptr, 0, 0) !=
            GPLUM_E_ARG)
            err_bad++;
        if (gplum_multiwalk_dispatch_index_mono(8, 1, e1, &ni, ij1, &one, is1, &zero,
                                                reinterpret_cast<const gplum_epj *>(epj.data()), nj, nullptr, 0,
                                                1) != GPLUM_E_ARG)
            err_bad++;
    }
    std::printf("fdps functors: rms rel acc err vs FP64 direct %.3e, neighbour-count mismatches %d, "
                "multiwalk mismatches %d, monopole multiwalk mismatches %d, index-vs-list mismatches %d, "
                "index argument checks failed %d\n", rms, bad_ngb, mw_bad, mono_bad, idx_bad, err_bad);
    if (rms > 1e-6 || bad_ngb != 0 || mw_bad != 0 || mono_bad != 0 || idx_bad != 0 || err_bad != 0) {
        std::printf("fdps functors: FAIL\n");
        return 1;
    }
    std::printf("fdps functors: OK\n");
    return 0;
}
