// radii.hip -- setCutoffRadii on the device: calcRandomVel (non-isotropic branch,
// src/particle.h:1319-1457) then setROutRSearch (src/particle.h:713-733).
//
// Every expression keeps the operation order of the reference (and of the oracle),
// so r_out / r_search come out bit-identical to the oracle's: the radial range is a
// max-reduction (exact in any order), the 32 bin sums run over each bin's particles in
// ascending id in a fixed blocked order (after a radix sort by (bin, id)), and the transcendental
// functions are the shared exact-operation versions (exp_exact_ops, cbrt_exact_ops).
#include <cfloat>

#include <hipcub/hipcub.hpp>

#include "system.hpp"

namespace gplum {

constexpr int RADII_NB = 32;
constexpr int RADII_ID_BITS = 58;

__global__ void k_radii_range(int n, const double *__restrict__ pos, double *__restrict__ part) {
    __shared__ double s_a[256], s_b[256];
    const int t = threadIdx.x;
    double a = 0., b = -DBL_MAX;
    for (int i = blockIdx.x * blockDim.x + t; i < n; i += gridDim.x * blockDim.x) {
        const double r2 = pos[3 * i] * pos[3 * i] + pos[3 * i + 1] * pos[3 * i + 1];
        a = smax(a, r2);
        b = smax(b, -r2);
    }
    s_a[t] = a;
    s_b[t] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            s_a[t] = smax(s_a[t], s_a[t + w]);
            s_b[t] = smax(s_b[t], s_b[t + w]);
        }
        __syncthreads();
    }
    if (t == 0) {
        part[2 * blockIdx.x] = s_a[0];
        part[2 * blockIdx.x + 1] = s_b[0];
    }
}

// range[0] = r_min, range[1] = dr, range[2] = 1/dr (one block of 256 threads)
__global__ __launch_bounds__(256) void k_radii_range_final(int nb, const double *__restrict__ part,
                                                           double *__restrict__ range) {
    __shared__ double s_a[256], s_b[256];
    const int t = threadIdx.x;
    double a = 0., b = -DBL_MAX;
    for (int k = t; k < nb; k += 256) {
        a = smax(a, part[2 * k]);
        b = smax(b, part[2 * k + 1]);
    }
    s_a[t] = a;
    s_b[t] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            s_a[t] = smax(s_a[t], s_a[t + w]);
            s_b[t] = smax(s_b[t], s_b[t + w]);
        }
        __syncthreads();
    }
    if (t != 0) return;
    double r_max = s_a[0], r_min = s_b[0];
    r_max = sqrt(r_max) * 1.01;
    r_min = sqrt(-r_min) * 0.99;
    const double dr = (r_max - r_min) / RADII_NB;
    range[0] = r_min;
    range[1] = dr;
    range[2] = 1. / dr;
}

__global__ void k_radii_bin(int n, const double *__restrict__ pos, const double *__restrict__ vel,
                            const int64_t *__restrict__ id, double m_sun, const double *__restrict__ range,
                            uint64_t *__restrict__ key, double *__restrict__ val, int *__restrict__ err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double r_min = range[0], drinv = range[2];
    const double x = pos[3 * i], y = pos[3 * i + 1];
    const double r2 = x * x + y * y;
    const double ri = sqrt(r2);
    int j = (int)((ri - r_min) * drinv);
    if (j == RADII_NB) j = RADII_NB - 1;
    const double vk = sqrt(m_sun / sqrt(x * x + y * y));
    double vkx = -y / ri, vky = x / ri, vkz = 0.;
    vkx *= vk; vky *= vk; vkz *= vk;
    const double ux = vel[3 * i] - vkx, uy = vel[3 * i + 1] - vky, uz = vel[3 * i + 2] - vkz;
    const int64_t idi = id[i];
    if (idi < 0 || idi >= (int64_t(1) << RADII_ID_BITS) || j < 0 || j >= RADII_NB) atomicOr(err, 1);
    key[i] = ((uint64_t)(j & 31) << RADII_ID_BITS) | ((uint64_t)idi & ((uint64_t(1) << RADII_ID_BITS) - 1));
    val[i] = ux * ux + uy * uy + uz * uz;
}

// Bin sums in a fixed blocked order: each bin's values in ascending id, summed in
// consecutive chunks of RADII_CHUNK (left to right), then the chunk sums left to
// right.  The oracle sums in the same blocked order, so v_disp agrees bit for bit.
constexpr int RADII_CHUNK = 64;

// one thread per bin: range of the bin in the (bin, id)-sorted keys and its chunks
__global__ void k_radii_binrange(int n, const uint64_t *__restrict__ key, int *__restrict__ range,
                                 int *__restrict__ cbase) {
    const int b = threadIdx.x;
    if (b < RADII_NB) {
        auto lower = [&](uint64_t k) {
            int lo = 0, hi = n;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (key[mid] < k) lo = mid + 1; else hi = mid;
            }
            return lo;
        };
        range[2 * b] = lower((uint64_t)b << RADII_ID_BITS);
        range[2 * b + 1] = lower((uint64_t)(b + 1) << RADII_ID_BITS);
    }
    __syncthreads();
    if (b == 0) {
        int c = 0;
        for (int q = 0; q < RADII_NB; q++) {
            cbase[q] = c;
            c += (range[2 * q + 1] - range[2 * q] + RADII_CHUNK - 1) / RADII_CHUNK;
        }
        cbase[RADII_NB] = c;
    }
}

__global__ void k_radii_chunksum(const int *__restrict__ range, const int *__restrict__ cbase,
                                 const double *__restrict__ val, double *__restrict__ csum) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cbase[RADII_NB]) return;
    int b = 0;
    while (b + 1 < RADII_NB && cbase[b + 1] <= c) b++;
    const int a0 = range[2 * b] + (c - cbase[b]) * RADII_CHUNK;
    const int a1 = min(a0 + RADII_CHUNK, range[2 * b + 1]);
    double s = 0.;
    for (int a = a0; a < a1; a++) s += val[a];
    csum[c] = s;
}

__global__ void k_radii_binfinal(const int *__restrict__ range, const int *__restrict__ cbase,
                                 const double *__restrict__ csum, double *__restrict__ vd, int *__restrict__ np) {
    const int b = threadIdx.x;
    if (b >= RADII_NB) return;
    // chunk sums added in ascending order; 32 loads in flight per batch (the adds
    // form one dependent chain, the loads need not)
    double s = 0.;
    int c = cbase[b];
    const int ce = cbase[b + 1];
    for (; c + 32 <= ce; c += 32) {
        double v[32];
#pragma unroll
        for (int k = 0; k < 32; k++) v[k] = csum[c + k];
#pragma unroll
        for (int k = 0; k < 32; k++) s += v[k];
    }
    for (; c < ce; c++) s += csum[c];
    const int cnt = range[2 * b + 1] - range[2 * b];
    np[b] = cnt;
    vd[b] = (cnt > 0) ? sqrt(s / cnt) : 0.;
}

struct RadiiParams {
    double m_sun, dt_tree, R_cut0, R_cut1, p_cut, r_cut_min, r_cut_max, R_search0, R_search1;
};

__global__ void k_radii_apply(int n, RadiiParams P, const double *__restrict__ range, const double *__restrict__ vdb,
                              const int *__restrict__ npb, const double *__restrict__ pos,
                              const double *__restrict__ vel, const double *__restrict__ mass,
                              double *__restrict__ v_disp, double *__restrict__ r_out,
                              double *__restrict__ r_out_inv, double *__restrict__ r_search) {
    __shared__ double s_vd[RADII_NB];
    __shared__ int s_np[RADII_NB];
    if (threadIdx.x < RADII_NB) { s_vd[threadIdx.x] = vdb[threadIdx.x]; s_np[threadIdx.x] = npb[threadIdx.x]; }
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double r_min = range[0], dr = range[1], drinv = range[2];
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    const double r2 = x * x + y * y;
    double vd = 0., ni = 0.;
    for (int b = 0; b < RADII_NB; b++) {
        const double ddr = r_min + (b + 0.5) * dr - sqrt(r2);
        const double expr = exp_exact_ops(-ddr * ddr * drinv * drinv);
        vd += s_vd[b] * expr;
        ni += (s_np[b] > 0) ? expr : 0.;
    }
    const double vdi = vd / ni;
    v_disp[i] = vdi;
    // getEccentricity / getSemimajorAxis2 / getRHill (particle.h:660-699)
    const double vx = vel[3 * i], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
    const double r = sqrt(x * x + y * y + z * z);
    const double rv = x * vx + y * vy + z * vz;
    const double ax = 1.0 / (2.0 / r - (vx * vx + vy * vy + vz * vz) / P.m_sun);
    const double ecccosu = 1. - r / ax;
    const double eccsinu2 = rv * rv / (P.m_sun * ax);
    const double ecc = sqrt(ecccosu * ecccosu + eccsinu2);
    const double ax2 = (ecc < 0.6) ? ax : sqrt(x * x + y * y + z * z);
    const double rHill = cbrt_exact_ops(mass[i] / (3. * P.m_sun)) * ax2;
    const double pc = (P.p_cut == 0.) ? 1. : pow(ax2, -P.p_cut);
    const double r_out_i = smax(P.R_cut0 * pc * rHill, P.R_cut1 * vdi * P.dt_tree);
    double ro = smax(r_out_i, P.r_cut_min);
    if (P.r_cut_max > 0.) ro = smin(P.r_cut_max, ro);
    r_out[i] = ro;
    r_out_inv[i] = 1. / ro;
    r_search[i] = P.R_search0 * ro + P.R_search1 * vdi * P.dt_tree;
}

int device_cutoff_radii(const gplum_params &p, ParticleSoA &A, int n, RadiiScratch &R, hipStream_t st) {
    if (n <= 0) return GPLUM_OK;
    const int nb = std::min(1024, ceil_div(n, 256));
    GPLUM_ENSURE(R.part, 2 * (size_t)nb);
    GPLUM_ENSURE(R.range, 4);
    GPLUM_ENSURE(R.key, n);
    GPLUM_ENSURE(R.key_sorted, n);
    GPLUM_ENSURE(R.val, n);
    GPLUM_ENSURE(R.val_sorted, n);
    GPLUM_ENSURE(R.vd, RADII_NB);
    GPLUM_ENSURE(R.np, RADII_NB);
    GPLUM_ENSURE(R.err, 1);
    GPLUM_HIP(hipMemsetAsync(R.err.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_radii_range, dim3(nb), dim3(256), 0, st, n, A.pos.p, R.part.p);
    hipLaunchKernelGGL(k_radii_range_final, dim3(1), dim3(256), 0, st, nb, R.part.p, R.range.p);
    hipLaunchKernelGGL(k_radii_bin, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, A.pos.p, A.vel.p, A.id.p, p.m_sun,
                       R.range.p, R.key.p, R.val.p, R.err.p);
    size_t bytes = 0;
    GPLUM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, R.key.p, R.key_sorted.p, R.val.p, R.val_sorted.p, n,
                                                  0, 64, st));
    GPLUM_ENSURE(R.tmp, bytes + 16);
    GPLUM_HIP(hipcub::DeviceRadixSort::SortPairs(R.tmp.p, bytes, R.key.p, R.key_sorted.p, R.val.p, R.val_sorted.p, n,
                                                  0, 64, st));
    GPLUM_ENSURE(R.brange, 2 * RADII_NB);
    GPLUM_ENSURE(R.cbase, RADII_NB + 1);
    const int max_chunks = ceil_div(n, RADII_CHUNK) + RADII_NB;
    GPLUM_ENSURE(R.csum, max_chunks);
    hipLaunchKernelGGL(k_radii_binrange, dim3(1), dim3(64), 0, st, n, R.key_sorted.p, R.brange.p, R.cbase.p);
    hipLaunchKernelGGL(k_radii_chunksum, dim3(ceil_div(max_chunks, 256)), dim3(256), 0, st, R.brange.p, R.cbase.p,
                       R.val_sorted.p, R.csum.p);
    hipLaunchKernelGGL(k_radii_binfinal, dim3(1), dim3(64), 0, st, R.brange.p, R.cbase.p, R.csum.p, R.vd.p, R.np.p);
    RadiiParams P{p.m_sun, p.dt_tree, p.R_cut0, p.R_cut1, p.p_cut, p.r_cut_min, p.r_cut_max, p.R_search0,
                  p.R_search1};
    hipLaunchKernelGGL(k_radii_apply, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, P, R.range.p, R.vd.p, R.np.p,
                       A.pos.p, A.vel.p, A.mass.p, A.v_disp.p, A.r_out.p, A.r_out_inv.p, A.r_search.p);
    GPLUM_HIP(hipGetLastError());
    // the flag (an id outside [0, 2^58), which set_particles rejects, or a bin index out of
    // range from a non-finite position) is read back without a wait and checked by
    // radii_check at the caller's next synchronisation
    GPLUM_HIP(R.err_host.ensure(sizeof(int)));
    GPLUM_HIP(hipMemcpyAsync(R.err_host.p, R.err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    R.err_pending = true;
    return GPLUM_OK;
}

int radii_check(RadiiScratch &R) {
    if (!R.err_pending) return GPLUM_OK;
    R.err_pending = false;
    if (*R.err_host.as<int>())
        return fail(GPLUM_E_ARG, "setCutoffRadii: particle ids must lie in [0, 2^58) and positions be finite");
    return GPLUM_OK;
}

}  // namespace gplum
