// soft_capi.hip -- PIKG-functor-shaped soft-force entry points (host arrays in/out).
//
// Each call uploads the caller's EPI/EPJ/SPJ/Force arrays into per-thread device
// buffers, runs one kernel over i-chunks of <= 256 particles (all relative to
// epi[0], as the PIKG kernel is) and downloads the accumulated Force array.  Calls
// from different host threads use different streams and buffers (FDPS calls the
// functors from many OpenMP threads, src/gravity_kernel.hpp:8-23).
#include <mutex>
#include <string>
#include <vector>

#include "soft_kernels.hpp"

namespace gplum {
namespace {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes) {
        if (bytes <= cap) return GPLUM_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = bytes < 4096 ? 4096 : bytes + bytes / 2;
        GPLUM_HIP(hipMalloc(&p, want));
        cap = want;
        return GPLUM_OK;
    }
};

struct CallCtx {
    hipStream_t st = nullptr;
    DevBuf epi, epj, spj, force, groups, acc, phi;
    std::vector<SoftGroup> h_groups;
    int init() {
        if (st) return GPLUM_OK;
        int count = 0;
        GPLUM_HIP(hipGetDeviceCount(&count));
        if (count <= 0) return fail(GPLUM_E_NODEVICE, "no HIP device");
        GPLUM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        return GPLUM_OK;
    }
};

thread_local CallCtx t_ctx;

int make_groups(CallCtx &c, int ni, int nj, int ns) {
    c.h_groups.clear();
    for (int b = 0; b < ni; b += SOFT_BLOCK) {
        SoftGroup g{};
        g.i_begin = b;
        g.ni = (ni - b) < SOFT_BLOCK ? (ni - b) : SOFT_BLOCK;
        g.j_begin = 0;
        g.nj = nj;
        g.jf_begin = 0;
        g.njf = 0;
        g.s_begin = 0;
        g.ns = ns;
        g.i_ref = 0;
        c.h_groups.push_back(g);
    }
    int rc = c.groups.ensure(sizeof(SoftGroup) * c.h_groups.size());
    if (rc) return rc;
    GPLUM_HIP(hipMemcpyAsync(c.groups.p, c.h_groups.data(), sizeof(SoftGroup) * c.h_groups.size(),
                             hipMemcpyHostToDevice, c.st));
    return GPLUM_OK;
}

template <class SP, class SPSRC, int MODE, bool F64>
int run_call(const gplum_epi *epi, int ni, const void *jp, size_t jsize, int nj,
             gplum_force *force, float eps2, double *acc64, double *phi64) {
    if (ni < 0 || nj < 0 || (ni > 0 && (!epi || !force)) || (nj > 0 && !jp))
        return fail(GPLUM_E_ARG, "null pointer or negative size");
    if (ni == 0) return GPLUM_OK;
    CallCtx &c = t_ctx;
    int rc = c.init();
    if (rc) return rc;
    if ((rc = c.epi.ensure(sizeof(gplum_epi) * ni))) return rc;
    if ((rc = c.epj.ensure(jsize * (size_t)(nj > 0 ? nj : 1)))) return rc;
    if ((rc = c.force.ensure(sizeof(gplum_force) * ni))) return rc;
    GPLUM_HIP(hipMemcpyAsync(c.epi.p, epi, sizeof(gplum_epi) * ni, hipMemcpyHostToDevice, c.st));
    if (nj > 0)
        GPLUM_HIP(hipMemcpyAsync(c.epj.p, jp, jsize * nj, hipMemcpyHostToDevice, c.st));
    GPLUM_HIP(hipMemcpyAsync(c.force.p, force, sizeof(gplum_force) * ni, hipMemcpyHostToDevice, c.st));
    const bool is_ep = (MODE & EP) != 0;
    if ((rc = make_groups(c, ni, is_ep ? nj : 0, is_ep ? 0 : nj))) return rc;
    const int ng = (int)c.h_groups.size();
    SrcEPI is{(const gplum_epi *)c.epi.p};
    CandBuf nocand{nullptr, nullptr, 0};
    if (!F64) {
        SinkForce out{(gplum_force *)c.force.p};
        if (is_ep) {
            launch_soft_f32<SrcEPI, SrcEPJ, SrcNone, SinkForce, MODE>(
                (const SoftGroup *)c.groups.p, ng, is, SrcEPJ{(const gplum_epj *)c.epj.p},
                SrcEPJ{(const gplum_epj *)c.epj.p}, SrcNone{},
                out, eps2, nocand, c.st);
        } else {
            launch_soft_f32<SrcEPI, SrcNone, SPSRC, SinkForce, MODE>(
                (const SoftGroup *)c.groups.p, ng, is, SrcNone{}, SrcNone{}, SPSRC{(const SP *)c.epj.p}, out,
                eps2, nocand, c.st);
        }
    } else {
        if ((rc = c.acc.ensure(sizeof(double) * 3 * ni))) return rc;
        if ((rc = c.phi.ensure(sizeof(double) * ni))) return rc;
        GPLUM_HIP(hipMemsetAsync(c.acc.p, 0, sizeof(double) * 3 * ni, c.st));
        GPLUM_HIP(hipMemsetAsync(c.phi.p, 0, sizeof(double) * ni, c.st));
        SinkF64 out{(double *)c.acc.p, (double *)c.phi.p, (gplum_force *)c.force.p};
        if (is_ep) {
            launch_soft_f64<SrcEPI, SrcEPJ, SrcNone, SinkF64, MODE>(
                (const SoftGroup *)c.groups.p, ng, is, SrcEPJ{(const gplum_epj *)c.epj.p},
                SrcEPJ{(const gplum_epj *)c.epj.p}, SrcNone{},
                out, (double)eps2, CandBuf{nullptr, nullptr, 0u}, c.st);
        } else {
            launch_soft_f64<SrcEPI, SrcNone, SPSRC, SinkF64, MODE>(
                (const SoftGroup *)c.groups.p, ng, is, SrcNone{}, SrcNone{}, SPSRC{(const SP *)c.epj.p}, out,
                (double)eps2, CandBuf{nullptr, nullptr, 0u}, c.st);
        }
        if (acc64)
            GPLUM_HIP(hipMemcpyAsync(acc64, c.acc.p, sizeof(double) * 3 * ni, hipMemcpyDeviceToHost, c.st));
        if (phi64)
            GPLUM_HIP(hipMemcpyAsync(phi64, c.phi.p, sizeof(double) * ni, hipMemcpyDeviceToHost, c.st));
    }
    GPLUM_HIP(hipGetLastError());
    GPLUM_HIP(hipMemcpyAsync(force, c.force.p, sizeof(gplum_force) * ni, hipMemcpyDeviceToHost, c.st));
    GPLUM_HIP(hipStreamSynchronize(c.st));
    return GPLUM_OK;
}

// ------------------------------------------------ no-allocation EP-EP variant
//
// The i-chunk table of the kernel (one SoftGroup per SOFT_BLOCK i-particles) is the
// only device memory the call needs beyond its inputs and outputs.  Here it lives in a
// fixed file-scope device array: each device holds its own copy until the process
// ends, every call on that device reuses it (calls that overlap on one device race for
// it), and the call allocates nothing.  Host code never passes the array's name to a
// kernel (in host code it is a stand-in, not a device address): the address comes
// from hipGetSymbolAddress, once per device, with that device current.
constexpr int NOALLOC_MAX_GROUPS = 4096;              // ni <= 4096 * 256 = 1,048,576
static __device__ __align__(256) SoftGroup scratch[NOALLOC_MAX_GROUPS];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
SoftGroup *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

int scratch_of_current_device(SoftGroup **out) {
    int dev = 0;
    GPLUM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(GPLUM_E_NODEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        GPLUM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(GPLUM_E_HIP, "hipGetSymbolAddress returned a null scratch address");
        g_scratch_ptr[dev] = static_cast<SoftGroup *>(p);
    }
    *out = g_scratch_ptr[dev];
    return GPLUM_OK;
}

// the i-chunk table of make_groups, written on the device (every entry is rewritten by
// each call, so nothing of an earlier call is read)
__global__ void k_noalloc_groups(SoftGroup *g, int ng, int ni, int nj) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ng) return;
    SoftGroup s{};
    s.i_begin = k * SOFT_BLOCK;
    s.ni = min(SOFT_BLOCK, ni - s.i_begin);
    s.j_begin = 0;
    s.nj = nj;
    s.jf_begin = 0;
    s.njf = 0;
    s.s_begin = 0;
    s.ns = 0;
    s.i_ref = 0;
    s.pad = 0;
    g[k] = s;
}

}  // namespace
}  // namespace gplum

using namespace gplum;

extern "C" {

int gplum_calc_force_epep(const gplum_epi *epi, int ni, const gplum_epj *epj, int nj,
                          gplum_force *force, float eps2) {
    return run_call<gplum_spj_quad, SrcSPQ, EP, false>(epi, ni, epj, sizeof(gplum_epj), nj, force,
                                                       eps2, nullptr, nullptr);
}

int gplum_calc_force_epsp_quad(const gplum_epi *epi, int ni, const gplum_spj_quad *spj, int nj,
                               gplum_force *force, float eps2) {
    return run_call<gplum_spj_quad, SrcSPQ, SPQ, false>(epi, ni, spj, sizeof(gplum_spj_quad), nj,
                                                        force, eps2, nullptr, nullptr);
}

int gplum_calc_force_epsp_mono(const gplum_epi *epi, int ni, const gplum_spj_mono *spj, int nj,
                               gplum_force *force, float eps2) {
    return run_call<gplum_spj_mono, SrcSPM, SPM, false>(epi, ni, spj, sizeof(gplum_spj_mono), nj,
                                                        force, eps2, nullptr, nullptr);
}

int gplum_calc_force_epep_f64(const gplum_epi *epi, int ni, const gplum_epj *epj, int nj,
                              gplum_force *force, float eps2, double *acc64, double *phi64) {
    return run_call<gplum_spj_quad, SrcSPQ, EP, true>(epi, ni, epj, sizeof(gplum_epj), nj, force,
                                                      eps2, acc64, phi64);
}

int gplum_calc_force_epsp_quad_f64(const gplum_epi *epi, int ni, const gplum_spj_quad *spj, int nj,
                                   gplum_force *force, float eps2, double *acc64, double *phi64) {
    return run_call<gplum_spj_quad, SrcSPQ, SPQ, true>(epi, ni, spj, sizeof(gplum_spj_quad), nj,
                                                       force, eps2, acc64, phi64);
}

int gplum_calc_force_epep_noalloc(const gplum_epi *d_epi, int ni, const gplum_epj *d_epj, int nj,
                                  gplum_force *d_force, float eps2, void *stream) {
    if (ni < 0 || nj < 0 || (ni > 0 && (!d_epi || !d_force)) || (nj > 0 && !d_epj))
        return fail(GPLUM_E_ARG, "null pointer or negative size");
    if (ni > NOALLOC_MAX_GROUPS * SOFT_BLOCK)
        return fail(GPLUM_E_ARG, "ni = " + std::to_string(ni) + " exceeds the no-allocation limit of " +
                                     std::to_string(NOALLOC_MAX_GROUPS * SOFT_BLOCK) + " i-particles");
    if (ni == 0) return GPLUM_OK;
    SoftGroup *groups = nullptr;
    int rc = scratch_of_current_device(&groups);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int ng = ceil_div(ni, SOFT_BLOCK);
    hipLaunchKernelGGL(k_noalloc_groups, dim3(ceil_div(ng, 256)), dim3(256), 0, st, groups, ng, ni, nj);
    launch_soft_f32<SrcEPI, SrcEPJ, SrcNone, SinkForce, EP>(groups, ng, SrcEPI{d_epi}, SrcEPJ{d_epj},
                                                             SrcEPJ{d_epj}, SrcNone{}, SinkForce{d_force}, eps2,
                                                             CandBuf{nullptr, nullptr, 0u}, st);
    GPLUM_HIP(hipGetLastError());
    return GPLUM_OK;
}

}  // extern "C"
