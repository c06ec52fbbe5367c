// multiwalk.hip -- FDPS calcForceAllAndWriteBackMultiWalk dispatch / retrieve pair.
//
// FDPS hands the user n_walk i-groups at a time, each with its EP and SP lists, via
//   S32 dispatch(S32 tag, S32 n_walk, const EPI** epi, const S32* n_epi,
//                const EPJ** epj, const S32* n_epj, const SPJ** spj, const S32* n_spj)
// and later collects the forces with
//   S32 retrieve(S32 tag, S32 n_walk, const S32* ni, Force** force).
// dispatch packs the walks into pinned staging buffers, uploads them and launches one
// EP-EP + EP-SP kernel over all walks on the tag's stream without waiting; retrieve
// waits for that tag and copies the forces out.  Forces start from the
// ForceGrav::clear() values and are assigned (not accumulated) to the caller's
// arrays.  Each walk's coordinates are relative to its first i-particle, as the PIKG
// kernels use epi[0] (gravity_kernel_epep.pikg:53-60).
#include <cstring>
#include <mutex>
#include <vector>

#include "soft_kernels.hpp"

namespace gplum {
namespace {

struct Pinned {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes) {
        if (bytes <= cap) return GPLUM_OK;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        size_t want = bytes + bytes / 2 + 4096;
        GPLUM_HIP(hipHostMalloc(&p, want, hipHostMallocDefault));
        cap = want;
        return GPLUM_OK;
    }
};
struct Dev {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes) {
        if (bytes <= cap) return GPLUM_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        size_t want = bytes + bytes / 2 + 4096;
        GPLUM_HIP(hipMalloc(&p, want));
        cap = want;
        return GPLUM_OK;
    }
};

struct TagCtx {
    hipStream_t st = nullptr;
    Pinned h_in, h_force;
    Dev d_in, d_force, d_groups;
    std::vector<int> i_off;      // per walk, into the packed i array
    int n_walk = 0, ni_total = 0;
    bool pending = false;
};

std::mutex g_mu;
std::vector<TagCtx *> g_tags;
float g_eps2 = 0.f;   // FP_t::eps2 as the PIKG functor receives it (gravity_kernel.hpp:18)

TagCtx *tag_ctx(int tag) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (tag < 0) return nullptr;
    if ((int)g_tags.size() <= tag) g_tags.resize(tag + 1, nullptr);
    if (!g_tags[tag]) g_tags[tag] = new TagCtx();
    return g_tags[tag];
}

inline size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace
}  // namespace gplum

using namespace gplum;

extern "C" int gplum_multiwalk_dispatch(int tag, int n_walk, const gplum_epi **epi, const int *n_epi,
                                        const gplum_epj **epj, const int *n_epj,
                                        const gplum_spj_quad **spj, const int *n_spj) {
    if (n_walk < 0 || (n_walk > 0 && (!epi || !n_epi || !epj || !n_epj || !spj || !n_spj)))
        return fail(GPLUM_E_ARG, "bad multiwalk arguments");
    TagCtx *c = tag_ctx(tag);
    if (!c) return fail(GPLUM_E_ARG, "negative tag");
    if (!c->st) GPLUM_HIP(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    if (c->pending) GPLUM_HIP(hipStreamSynchronize(c->st));
    long long ni = 0, nj = 0, ns = 0;
    for (int w = 0; w < n_walk; w++) {
        if (n_epi[w] < 0 || n_epj[w] < 0 || n_spj[w] < 0) return fail(GPLUM_E_ARG, "negative list size");
        ni += n_epi[w];
        nj += n_epj[w];
        ns += n_spj[w];
    }
    // packed layout: [epi | epj | spj | groups]
    std::vector<SoftGroup> groups;
    c->i_off.assign(n_walk + 1, 0);
    long long io = 0, jo = 0, so = 0;
    for (int w = 0; w < n_walk; w++) {
        c->i_off[w] = (int)io;
        for (int b = 0; b < n_epi[w]; b += SOFT_BLOCK) {
            SoftGroup g{};
            g.i_begin = (int)(io + b);
            g.ni = std::min(SOFT_BLOCK, n_epi[w] - b);
            g.j_begin = (int)jo;
            g.nj = n_epj[w];
            g.jf_begin = 0;
            g.njf = 0;
            g.s_begin = (int)so;
            g.ns = n_spj[w];
            g.i_ref = (int)io;
            groups.push_back(g);
        }
        io += n_epi[w];
        jo += n_epj[w];
        so += n_spj[w];
    }
    c->i_off[n_walk] = (int)io;
    c->n_walk = n_walk;
    c->ni_total = (int)ni;
    const size_t b_epi = align_up(sizeof(gplum_epi) * ni), b_epj = align_up(sizeof(gplum_epj) * nj);
    const size_t b_spj = align_up(sizeof(gplum_spj_quad) * ns), b_grp = align_up(sizeof(SoftGroup) * groups.size());
    const size_t total = b_epi + b_epj + b_spj + b_grp;
    int rc;
    if ((rc = c->h_in.ensure(total)) || (rc = c->d_in.ensure(total))) return rc;
    if ((rc = c->h_force.ensure(sizeof(gplum_force) * (ni + 1))) || (rc = c->d_force.ensure(sizeof(gplum_force) * (ni + 1))))
        return rc;
    char *h = (char *)c->h_in.p;
    size_t o = 0;
    for (int w = 0; w < n_walk; w++) {
        std::memcpy(h + o, epi[w], sizeof(gplum_epi) * n_epi[w]);
        o += sizeof(gplum_epi) * n_epi[w];
    }
    o = b_epi;
    for (int w = 0; w < n_walk; w++) {
        std::memcpy(h + o, epj[w], sizeof(gplum_epj) * n_epj[w]);
        o += sizeof(gplum_epj) * n_epj[w];
    }
    o = b_epi + b_epj;
    for (int w = 0; w < n_walk; w++) {
        std::memcpy(h + o, spj[w], sizeof(gplum_spj_quad) * n_spj[w]);
        o += sizeof(gplum_spj_quad) * n_spj[w];
    }
    std::memcpy(h + b_epi + b_epj + b_spj, groups.data(), sizeof(SoftGroup) * groups.size());
    gplum_force *hf = (gplum_force *)c->h_force.p;
    for (long long i = 0; i < ni; i++) {
        hf[i].acc[0] = hf[i].acc[1] = hf[i].acc[2] = 0.f;
        hf[i].phi = 0.f;
        hf[i].neighbor.number = 0;
        hf[i].neighbor.rank = 0;
        hf[i].neighbor.id_max = ID_MAX_CLEAR;
        hf[i].neighbor.id_min = ID_MIN_CLEAR;
    }
    GPLUM_HIP(hipMemcpyAsync(c->d_in.p, c->h_in.p, total, hipMemcpyHostToDevice, c->st));
    GPLUM_HIP(hipMemcpyAsync(c->d_force.p, c->h_force.p, sizeof(gplum_force) * ni, hipMemcpyHostToDevice, c->st));
    char *d = (char *)c->d_in.p;
    SrcEPI is{(const gplum_epi *)d};
    SrcEPJ js{(const gplum_epj *)(d + b_epi)};
    SrcSPQ ss{(const gplum_spj_quad *)(d + b_epi + b_epj)};
    SinkForce out{(gplum_force *)c->d_force.p};
    CandBuf nocand{nullptr, nullptr, 0};
    launch_soft_f32<SrcEPI, SrcEPJ, SrcSPQ, SinkForce, EP | SPQ>((const SoftGroup *)(d + b_epi + b_epj + b_spj),
                                                                 (int)groups.size(), is, js, js, ss, out, g_eps2,
                                                                 nocand, c->st);
    GPLUM_HIP(hipGetLastError());
    GPLUM_HIP(hipMemcpyAsync(c->h_force.p, c->d_force.p, sizeof(gplum_force) * ni, hipMemcpyDeviceToHost, c->st));
    c->pending = true;
    return GPLUM_OK;
}

extern "C" int gplum_multiwalk_set_eps2(float eps2) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_eps2 = eps2;
    return GPLUM_OK;
}

extern "C" int gplum_multiwalk_retrieve(int tag, int n_walk, const int *ni, gplum_force **force) {
    TagCtx *c = tag_ctx(tag);
    if (!c || !c->pending) return fail(GPLUM_E_STATE, "retrieve without dispatch");
    if (n_walk != c->n_walk) return fail(GPLUM_E_ARG, "n_walk differs from dispatch");
    GPLUM_HIP(hipStreamSynchronize(c->st));
    const gplum_force *hf = (const gplum_force *)c->h_force.p;
    for (int w = 0; w < n_walk; w++) {
        const int cnt = c->i_off[w + 1] - c->i_off[w];
        if (ni && ni[w] != cnt) return fail(GPLUM_E_ARG, "ni differs from dispatch");
        std::memcpy(force[w], hf + c->i_off[w], sizeof(gplum_force) * cnt);
    }
    c->pending = false;
    return GPLUM_OK;
}
