// multiwalk.hip -- FDPS calcForceAllAndWriteBackMultiWalk[Index] dispatch / retrieve pair.
//
// FDPS hands the user n_walk i-groups at a time, each with its EP and SP lists, via
//   S32 dispatch(S32 tag, S32 n_walk, const EPI** epi, const S32* n_epi,
//                const EPJ** epj, const S32* n_epj, const SPJ** spj, const S32* n_spj)
// or, in the Index variant, with per-walk index lists into the tree's whole EPJ/SPJ
// arrays, which are sent only when send_flag is set:
//   S32 dispatch(S32 tag, S32 n_walk, const EPI** epi, const S32* n_epi,
//                const S32** id_epj, const S32* n_epj, const S32** id_spj, const S32* n_spj,
//                const EPJ* epj, const S32 n_epj_tot, const SPJ* spj, const S32 n_spj_tot,
//                const bool send_flag)
// and later collects the forces with
//   S32 retrieve(S32 tag, S32 n_walk, const S32* ni, Force** force).
// dispatch packs the walks into pinned staging buffers, uploads them and launches one
// EP-EP + EP-SP kernel over all walks on the tag's stream without waiting; retrieve
// waits for that tag and copies the forces out.  Forces start from the
// ForceGrav::clear() values and are assigned (not accumulated) to the caller's
// arrays.  Each walk's coordinates are relative to its first i-particle, as the PIKG
// kernels use epi[0] (gravity_kernel_epep.pikg:53-60).  SPJ is the quadrupole or the
// monopole super-particle (USE_QUAD on or off).
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
    Dev d_in, d_force;
    std::vector<int> i_off;      // per walk, into the packed i array
    int n_walk = 0, ni_total = 0;
    bool pending = false;
};

// The Index variant's whole EPJ / SPJ arrays, shared by the tags: sent once per force
// call with send_flag (a synchronous copy, so every later launch of any tag sees them;
// FDPS has retrieved every tag of the previous force call before it sends again).
struct IndexArrays {
    Dev epj, spj;
    int n_epj = 0, n_spj = 0, spj_size = 0;
    bool valid = false;
};

std::mutex g_mu;
std::vector<TagCtx *> g_tags;
IndexArrays g_index;
float g_eps2 = 0.f;   // FP_t::eps2 as the PIKG functor receives it (gravity_kernel.hpp:18)

TagCtx *tag_ctx(int tag) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (tag < 0) return nullptr;
    if ((int)g_tags.size() <= tag) g_tags.resize(tag + 1, nullptr);
    if (!g_tags[tag]) g_tags[tag] = new TagCtx();
    return g_tags[tag];
}

inline size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

int begin_dispatch(int tag, int n_walk, const void *epi, const int *n_epi, const void *jl, const int *n_epj,
                   const void *sl, const int *n_spj, TagCtx **out, long long sums[3]) {
    if (n_walk < 0 || (n_walk > 0 && (!epi || !n_epi || !jl || !n_epj || !sl || !n_spj)))
        return fail(GPLUM_E_ARG, "bad multiwalk arguments");
    TagCtx *c = tag_ctx(tag);
    if (!c) return fail(GPLUM_E_ARG, "negative tag");
    if (!c->st) GPLUM_HIP(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    if (c->pending) GPLUM_HIP(hipStreamSynchronize(c->st));
    sums[0] = sums[1] = sums[2] = 0;
    for (int w = 0; w < n_walk; w++) {
        if (n_epi[w] < 0 || n_epj[w] < 0 || n_spj[w] < 0) return fail(GPLUM_E_ARG, "negative list size");
        sums[0] += n_epi[w];
        sums[1] += n_epj[w];
        sums[2] += n_spj[w];
    }
    if (sums[0] > INT32_MAX || sums[1] > INT32_MAX || sums[2] > INT32_MAX)
        return fail(GPLUM_E_OVERFLOW, "multiwalk list totals exceed 2^31");
    *out = c;
    return GPLUM_OK;
}

// i-chunks of SOFT_BLOCK per walk; j and s entries of walk w start at the running totals
std::vector<SoftGroup> walk_groups(TagCtx *c, int n_walk, const int *n_epi, const int *n_epj, const int *n_spj) {
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
    c->ni_total = (int)io;
    return groups;
}

// packs arrays of `size` bytes per element, walk after walk, at h + o
size_t pack(char *h, size_t o, int n_walk, const void *const *src, const int *cnt, size_t size) {
    for (int w = 0; w < n_walk; w++) {
        std::memcpy(h + o, src[w], size * (size_t)cnt[w]);
        o += size * (size_t)cnt[w];
    }
    return o;
}

int clear_forces(TagCtx *c, long long ni) {
    int rc;
    if ((rc = c->h_force.ensure(sizeof(gplum_force) * (ni + 1))) || (rc = c->d_force.ensure(sizeof(gplum_force) * (ni + 1))))
        return rc;
    gplum_force *hf = (gplum_force *)c->h_force.p;
    for (long long i = 0; i < ni; i++) {
        hf[i].acc[0] = hf[i].acc[1] = hf[i].acc[2] = 0.f;
        hf[i].phi = 0.f;
        hf[i].neighbor.number = 0;
        hf[i].neighbor.rank = 0;
        hf[i].neighbor.id_max = ID_MAX_CLEAR;
        hf[i].neighbor.id_min = ID_MIN_CLEAR;
    }
    GPLUM_HIP(hipMemcpyAsync(c->d_force.p, c->h_force.p, sizeof(gplum_force) * ni, hipMemcpyHostToDevice, c->st));
    return GPLUM_OK;
}

template <class JS, class SS, int SPMODE>
int launch_and_return(TagCtx *c, const SoftGroup *d_groups, int ng, SrcEPI is, JS js, SS ss) {
    SinkForce out{(gplum_force *)c->d_force.p};
    CandBuf nocand{nullptr, nullptr, 0};
    launch_soft_f32<SrcEPI, JS, SS, SinkForce, EP | SPMODE>(d_groups, ng, is, js, js, ss, out, g_eps2, nocand, c->st);
    GPLUM_HIP(hipGetLastError());
    GPLUM_HIP(hipMemcpyAsync(c->h_force.p, c->d_force.p, sizeof(gplum_force) * c->ni_total, hipMemcpyDeviceToHost,
                             c->st));
    c->pending = true;
    return GPLUM_OK;
}

// the list variant: per-walk EPJ and SPJ arrays packed after the EPI arrays
template <class SPT, class SPSRC, int SPMODE>
int dispatch_lists(int tag, int n_walk, const gplum_epi **epi, const int *n_epi, const gplum_epj **epj,
                   const int *n_epj, const SPT **spj, const int *n_spj) {
    TagCtx *c = nullptr;
    long long sums[3];
    int rc = begin_dispatch(tag, n_walk, epi, n_epi, epj, n_epj, spj, n_spj, &c, sums);
    if (rc) return rc;
    const long long ni = sums[0], nj = sums[1], ns = sums[2];
    std::vector<SoftGroup> groups = walk_groups(c, n_walk, n_epi, n_epj, n_spj);
    // packed layout: [epi | epj | spj | groups]
    const size_t b_epi = align_up(sizeof(gplum_epi) * ni), b_epj = align_up(sizeof(gplum_epj) * nj);
    const size_t b_spj = align_up(sizeof(SPT) * ns), b_grp = align_up(sizeof(SoftGroup) * groups.size());
    const size_t total = b_epi + b_epj + b_spj + b_grp;
    if ((rc = c->h_in.ensure(total)) || (rc = c->d_in.ensure(total))) return rc;
    char *h = (char *)c->h_in.p;
    pack(h, 0, n_walk, (const void *const *)epi, n_epi, sizeof(gplum_epi));
    pack(h, b_epi, n_walk, (const void *const *)epj, n_epj, sizeof(gplum_epj));
    pack(h, b_epi + b_epj, n_walk, (const void *const *)spj, n_spj, sizeof(SPT));
    std::memcpy(h + b_epi + b_epj + b_spj, groups.data(), sizeof(SoftGroup) * groups.size());
    if ((rc = clear_forces(c, ni))) return rc;
    GPLUM_HIP(hipMemcpyAsync(c->d_in.p, c->h_in.p, total, hipMemcpyHostToDevice, c->st));
    char *d = (char *)c->d_in.p;
    return launch_and_return<SrcEPJ, SPSRC, SPMODE>(
        c, (const SoftGroup *)(d + b_epi + b_epj + b_spj), (int)groups.size(), SrcEPI{(const gplum_epi *)d},
        SrcEPJ{(const gplum_epj *)(d + b_epi)}, SPSRC{(const SPT *)(d + b_epi + b_epj)});
}

// the Index variant: per-walk index lists into the whole arrays (uploaded when sent)
template <class SPT, int SPMODE>
int dispatch_index(int tag, int n_walk, const gplum_epi **epi, const int *n_epi, const int32_t **id_epj,
                   const int *n_epj, const int32_t **id_spj, const int *n_spj, const gplum_epj *epj,
                   int n_epj_tot, const SPT *spj, int n_spj_tot, int send_flag) {
    TagCtx *c = nullptr;
    long long sums[3];
    int rc = begin_dispatch(tag, n_walk, epi, n_epi, id_epj, n_epj, id_spj, n_spj, &c, sums);
    if (rc) return rc;
    const long long ni = sums[0], nj = sums[1], ns = sums[2];
    {
        std::lock_guard<std::mutex> lk(g_mu);
        IndexArrays &A = g_index;
        if (send_flag) {
            if (n_epj_tot < 0 || n_spj_tot < 0 || (n_epj_tot > 0 && !epj) || (n_spj_tot > 0 && !spj))
                return fail(GPLUM_E_ARG, "bad whole EPJ/SPJ arrays");
            if ((rc = A.epj.ensure(sizeof(gplum_epj) * ((size_t)n_epj_tot + 1))) ||
                (rc = A.spj.ensure(sizeof(SPT) * ((size_t)n_spj_tot + 1))))
                return rc;
            if (n_epj_tot > 0) GPLUM_HIP(hipMemcpy(A.epj.p, epj, sizeof(gplum_epj) * n_epj_tot, hipMemcpyHostToDevice));
            if (n_spj_tot > 0) GPLUM_HIP(hipMemcpy(A.spj.p, spj, sizeof(SPT) * n_spj_tot, hipMemcpyHostToDevice));
            A.n_epj = n_epj_tot;
            A.n_spj = n_spj_tot;
            A.spj_size = (int)sizeof(SPT);
            A.valid = true;
        } else {
            if (!A.valid) return fail(GPLUM_E_STATE, "index dispatch before the whole arrays were sent");
            if (A.spj_size != (int)sizeof(SPT)) return fail(GPLUM_E_ARG, "SPJ type differs from the sent array");
        }
    }
    // the indices must lie in the whole arrays: checked on the host, as FDPS would assert
    for (int w = 0; w < n_walk; w++) {
        for (int k = 0; k < n_epj[w]; k++)
            if (id_epj[w][k] < 0 || id_epj[w][k] >= g_index.n_epj) return fail(GPLUM_E_ARG, "EPJ index out of range");
        for (int k = 0; k < n_spj[w]; k++)
            if (id_spj[w][k] < 0 || id_spj[w][k] >= g_index.n_spj) return fail(GPLUM_E_ARG, "SPJ index out of range");
    }
    std::vector<SoftGroup> groups = walk_groups(c, n_walk, n_epi, n_epj, n_spj);
    // packed layout: [epi | id_epj | id_spj | groups]
    const size_t b_epi = align_up(sizeof(gplum_epi) * ni), b_ij = align_up(sizeof(int32_t) * nj);
    const size_t b_is = align_up(sizeof(int32_t) * ns), b_grp = align_up(sizeof(SoftGroup) * groups.size());
    const size_t total = b_epi + b_ij + b_is + b_grp;
    if ((rc = c->h_in.ensure(total)) || (rc = c->d_in.ensure(total))) return rc;
    char *h = (char *)c->h_in.p;
    pack(h, 0, n_walk, (const void *const *)epi, n_epi, sizeof(gplum_epi));
    pack(h, b_epi, n_walk, (const void *const *)id_epj, n_epj, sizeof(int32_t));
    pack(h, b_epi + b_ij, n_walk, (const void *const *)id_spj, n_spj, sizeof(int32_t));
    std::memcpy(h + b_epi + b_ij + b_is, groups.data(), sizeof(SoftGroup) * groups.size());
    if ((rc = clear_forces(c, ni))) return rc;
    GPLUM_HIP(hipMemcpyAsync(c->d_in.p, c->h_in.p, total, hipMemcpyHostToDevice, c->st));
    char *d = (char *)c->d_in.p;
    return launch_and_return<SrcEPJIdx, SrcSPIdx<SPT>, SPMODE>(
        c, (const SoftGroup *)(d + b_epi + b_ij + b_is), (int)groups.size(), SrcEPI{(const gplum_epi *)d},
        SrcEPJIdx{(const gplum_epj *)g_index.epj.p, (const int32_t *)(d + b_epi)},
        SrcSPIdx<SPT>{(const SPT *)g_index.spj.p, (const int32_t *)(d + b_epi + b_ij)});
}

}  // namespace
}  // namespace gplum

using namespace gplum;

extern "C" int gplum_multiwalk_dispatch(int tag, int n_walk, const gplum_epi **epi, const int *n_epi,
                                        const gplum_epj **epj, const int *n_epj,
                                        const gplum_spj_quad **spj, const int *n_spj) {
    return dispatch_lists<gplum_spj_quad, SrcSPQ, SPQ>(tag, n_walk, epi, n_epi, epj, n_epj, spj, n_spj);
}

extern "C" int gplum_multiwalk_dispatch_mono(int tag, int n_walk, const gplum_epi **epi, const int *n_epi,
                                             const gplum_epj **epj, const int *n_epj,
                                             const gplum_spj_mono **spj, const int *n_spj) {
    return dispatch_lists<gplum_spj_mono, SrcSPM, SPM>(tag, n_walk, epi, n_epi, epj, n_epj, spj, n_spj);
}

extern "C" int gplum_multiwalk_dispatch_index(int tag, int n_walk, const gplum_epi **epi, const int *n_epi,
                                              const int32_t **id_epj, const int *n_epj, const int32_t **id_spj,
                                              const int *n_spj, const gplum_epj *epj, int n_epj_tot,
                                              const gplum_spj_quad *spj, int n_spj_tot, int send_flag) {
    return dispatch_index<gplum_spj_quad, SPQ>(tag, n_walk, epi, n_epi, id_epj, n_epj, id_spj, n_spj, epj, n_epj_tot,
                                               spj, n_spj_tot, send_flag);
}

extern "C" int gplum_multiwalk_dispatch_index_mono(int tag, int n_walk, const gplum_epi **epi, const int *n_epi,
                                                   const int32_t **id_epj, const int *n_epj,
                                                   const int32_t **id_spj, const int *n_spj, const gplum_epj *epj,
                                                   int n_epj_tot, const gplum_spj_mono *spj, int n_spj_tot,
                                                   int send_flag) {
    return dispatch_index<gplum_spj_mono, SPM>(tag, n_walk, epi, n_epi, id_epj, n_epj, id_spj, n_spj, epj, n_epj_tot,
                                               spj, n_spj_tot, send_flag);
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
