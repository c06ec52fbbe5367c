// mlp.hip -- neighbour-count forecaster: a 4-64-64-1 MLP (tanh, softplus Poisson rate)
// trained online with hand-written forward / backward / reduction / Adam kernels.
//
// Parameter layout (float32, 4545 values): W1[64][4], b1[64], W2[64][64], b2[64],
// w3[64], b3.  Loss: mean Poisson NLL  lambda - y log(lambda + 1e-6).
// Forward/backward: one thread per sample, parameters read with wave-uniform addresses
// (scalar loads, no per-lane traffic).  Gradient: per-sample activations and deltas are stored
// sample-major ([n][64]); one lane per parameter sums its products over a chunk of
// samples in FP64 (lanes of a wave read one contiguous 256-byte row per sample), and a
// second pass adds the chunk partials in chunk order, so the gradient is deterministic.
#include <vector>

#include "mlp.hpp"

namespace gplum {

constexpr int OFF_W1 = 0, OFF_B1 = 256, OFF_W2 = 320, OFF_B2 = 4416, OFF_W3 = 4480, OFF_B3 = 4544;
static_assert(OFF_B3 + 1 == MLP_NPARAM, "parameter layout");
constexpr float MLP_EPS = 1e-6f;

__device__ inline float softplus_f(float z) { return z > 20.f ? z : log1pf(expf(z)); }
__device__ inline float sigmoid_f(float z) { return 1.f / (1.f + expf(-z)); }

// per-sample forward pass; h1, h2 in registers.  The weights are read with
// wave-uniform addresses straight from global memory, which the compiler turns into
// scalar loads: every lane uses the same weight, so no LDS traffic per FMA.
__device__ inline float mlp_eval(const float *__restrict__ s, const float x[MLP_NIN], float h1[MLP_NH],
                                 float h2[MLP_NH], float *zout) {
    // every loop fully unrolled: h1/h2 are indexed by compile-time constants and stay
    // in registers (a runtime index would put them in scratch memory)
#pragma unroll
    for (int p = 0; p < MLP_NH; p++) {
        float a = s[OFF_B1 + p];
#pragma unroll
        for (int q = 0; q < MLP_NIN; q++) a = __builtin_fmaf(s[OFF_W1 + p * MLP_NIN + q], x[q], a);
        h1[p] = tanhf(a);
    }
#pragma unroll
    for (int p = 0; p < MLP_NH; p++) {
        float a = s[OFF_B2 + p];
#pragma unroll
        for (int q = 0; q < MLP_NH; q++) a = __builtin_fmaf(s[OFF_W2 + p * MLP_NH + q], h1[q], a);
        h2[p] = tanhf(a);
    }
    float z = s[OFF_B3];
#pragma unroll
    for (int p = 0; p < MLP_NH; p++) z = __builtin_fmaf(s[OFF_W3 + p], h2[p], z);
    *zout = z;
    return softplus_f(z);
}

__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ theta, const float *__restrict__ x,
                                                     int n, float *__restrict__ lam) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float xi[MLP_NIN], h1[MLP_NH], h2[MLP_NH], z;
#pragma unroll
    for (int q = 0; q < MLP_NIN; q++) xi[q] = x[(size_t)q * n + i];
    lam[i] = mlp_eval(theta, xi, h1, h2, &z);
}

// per-sample backward: stores h1, h2, d1, d2 (sample-major, [n][64]) and dz[n]
__global__ __launch_bounds__(256) void k_mlp_backward(const float *__restrict__ theta, const float *__restrict__ x,
                                                      const float *__restrict__ y, int n, float *__restrict__ H1,
                                                      float *__restrict__ H2, float *__restrict__ D1,
                                                      float *__restrict__ D2, float *__restrict__ DZ,
                                                      double *__restrict__ loss_part) {
    __shared__ double sl[256];
    const float *__restrict__ s = theta;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double li = 0.;
    if (i < n) {
        float xi[MLP_NIN], h1[MLP_NH], h2[MLP_NH], z;
#pragma unroll
        for (int q = 0; q < MLP_NIN; q++) xi[q] = x[(size_t)q * n + i];
        const float lam = mlp_eval(s, xi, h1, h2, &z);
        const float yi = y[i];
        li = (double)lam - (double)yi * log((double)lam + (double)MLP_EPS);
        const float dz = (1.f - yi / (lam + MLP_EPS)) * sigmoid_f(z) / (float)n;
        DZ[i] = dz;
        float d2[MLP_NH];
#pragma unroll
        for (int p = 0; p < MLP_NH; p++) {
            d2[p] = dz * s[OFF_W3 + p] * (1.f - h2[p] * h2[p]);
            H1[(size_t)i * MLP_NH + p] = h1[p];
            H2[(size_t)i * MLP_NH + p] = h2[p];
            D2[(size_t)i * MLP_NH + p] = d2[p];
        }
#pragma unroll
        for (int q = 0; q < MLP_NH; q++) {
            float a = 0.f;
#pragma unroll
            for (int p = 0; p < MLP_NH; p++) a = __builtin_fmaf(s[OFF_W2 + p * MLP_NH + q], d2[p], a);
            D1[(size_t)i * MLP_NH + q] = a * (1.f - h1[q] * h1[q]);
        }
    }
    sl[threadIdx.x] = li;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sl[threadIdx.x] += sl[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = sl[0];
}

constexpr int MLP_CHUNK = 128;    // samples per partial sum (fixed: the sum order)

// grid (ceil(NPARAM/64), nchunk), one wave per block: lane = parameter, y = sample chunk.
// x is feature-major [4][n] (the kernels' input layout), the rest sample-major.
__global__ __launch_bounds__(64) void k_mlp_grad_part(const float *__restrict__ x, int n,
                                                      const float *__restrict__ H1, const float *__restrict__ H2,
                                                      const float *__restrict__ D1, const float *__restrict__ D2,
                                                      const float *__restrict__ DZ, double *__restrict__ part) {
    const int k = blockIdx.x * 64 + threadIdx.x;
    if (k >= MLP_NPARAM) return;
    const int s0 = blockIdx.y * MLP_CHUNK, s1 = min(n, s0 + MLP_CHUNK);
    double acc = 0.;
    if (k < OFF_B1) {
        const int p = k / MLP_NIN, q = k % MLP_NIN;
#pragma unroll 4
        for (int s = s0; s < s1; s++) acc += (double)D1[(size_t)s * MLP_NH + p] * (double)x[(size_t)q * n + s];
    } else if (k < OFF_W2) {
        const int p = k - OFF_B1;
#pragma unroll 4
        for (int s = s0; s < s1; s++) acc += (double)D1[(size_t)s * MLP_NH + p];
    } else if (k < OFF_B2) {
        const int p = (k - OFF_W2) / MLP_NH, q = (k - OFF_W2) % MLP_NH;
#pragma unroll 4
        for (int s = s0; s < s1; s++) acc += (double)D2[(size_t)s * MLP_NH + p] * (double)H1[(size_t)s * MLP_NH + q];
    } else if (k < OFF_W3) {
        const int p = k - OFF_B2;
#pragma unroll 4
        for (int s = s0; s < s1; s++) acc += (double)D2[(size_t)s * MLP_NH + p];
    } else if (k < OFF_B3) {
        const int p = k - OFF_W3;
#pragma unroll 4
        for (int s = s0; s < s1; s++) acc += (double)DZ[s] * (double)H2[(size_t)s * MLP_NH + p];
    } else {
#pragma unroll 4
        for (int s = s0; s < s1; s++) acc += (double)DZ[s];
    }
    part[(size_t)blockIdx.y * MLP_NPARAM + k] = acc;
}

__global__ void k_mlp_grad_sum(int nchunk, const double *__restrict__ part, float *__restrict__ grad) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= MLP_NPARAM) return;
    double acc = 0.;
    for (int c = 0; c < nchunk; c++) acc += part[(size_t)c * MLP_NPARAM + k];
    grad[k] = (float)acc;
}

__global__ void k_mlp_adam(float *__restrict__ theta, const float *__restrict__ grad, float *__restrict__ m,
                           float *__restrict__ v, float lr, float b1, float b2, float c1, float c2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= MLP_NPARAM) return;
    const float g = grad[k];
    const float mk = b1 * m[k] + (1.f - b1) * g;
    const float vk = b2 * v[k] + (1.f - b2) * g * g;
    m[k] = mk;
    v[k] = vk;
    theta[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + 1e-8f);
}

// ---------------------------------------------- prediction on the matrix cores
//
// The rates that size the candidate buffers (fc_plan) need no more than a few digits:
// the 64x64 hidden layer runs on MFMA (v_mfma_f32_16x16x16bf16_1k: bf16 operands,
// FP32 accumulation) and tanh comes from the hardware exp and reciprocal.  A wave
// takes 16 samples per tile: lane l evaluates layer 1 for sample l % 16 and the
// sixteen hidden units 16 s + 4 (l / 16) + j (s, j = 0..3), which are exactly its A
// operands (A[i][k] = h1 of sample i, unit k) of the four K-steps; W2 is held as the B
// operands (B[k][n] = W2[n][k]) for the four 16-column output tiles; the accumulators
// (D[i][n], i = 4 (l / 16) + j, n = 16 t + l % 16) start at b2, and the output layer
// sums w3 h2 over the 16 lanes of a row.  Agrees with k_mlp_forward to bf16 precision.
typedef short mlp_v4s __attribute__((ext_vector_type(4)));
typedef float mlp_v4f __attribute__((ext_vector_type(4)));

__device__ inline short bf16_rn(float x) {                // round to nearest even
    const unsigned u = __float_as_uint(x);
    return (short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline float tanh_fast(float a) {               // 1 - 2 / (1 + e^(2a))
    return 1.f - 2.f * __builtin_amdgcn_rcpf(1.f + __expf(2.f * a));
}

constexpr int MLP_PRED_BLOCK = 256;

__global__ __launch_bounds__(MLP_PRED_BLOCK) void k_mlp_predict(const float *__restrict__ s,
                                                                const float *__restrict__ x, int n,
                                                                float *__restrict__ lam) {
    const int l = (int)threadIdx.x & 63;
    const int r = l >> 4, c = l & 15;
    // this lane's layer-1 units, W2 as B operands, b2 / w3 of its output columns
    float w1[16][MLP_NIN], b1[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
        const int k = 16 * (u >> 2) + 4 * r + (u & 3);
        b1[u] = s[OFF_B1 + k];
#pragma unroll
        for (int q = 0; q < MLP_NIN; q++) w1[u][q] = s[OFF_W1 + k * MLP_NIN + q];
    }
    mlp_v4s bw[4][4];
    float b2[4], w3[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int nn = 16 * t + c;
        b2[t] = s[OFF_B2 + nn];
        w3[t] = s[OFF_W3 + nn];
#pragma unroll
        for (int ks = 0; ks < 4; ks++)
#pragma unroll
            for (int j = 0; j < 4; j++) bw[t][ks][j] = bf16_rn(s[OFF_W2 + nn * MLP_NH + 16 * ks + 4 * r + j]);
    }
    const float b3 = s[OFF_B3];
    const int waves = (int)(gridDim.x * (MLP_PRED_BLOCK / 64));
    const int tiles = (n + 15) / 16;
    for (int tile = (int)blockIdx.x * (MLP_PRED_BLOCK / 64) + ((int)threadIdx.x >> 6); tile < tiles;
         tile += waves) {
        const int i = 16 * tile + c;
        float xi[MLP_NIN];
#pragma unroll
        for (int q = 0; q < MLP_NIN; q++) xi[q] = (i < n) ? x[(size_t)q * n + i] : 0.f;
        mlp_v4s a[4];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            float z = b1[u];
#pragma unroll
            for (int q = 0; q < MLP_NIN; q++) z = __builtin_fmaf(w1[u][q], xi[q], z);
            a[u >> 2][u & 3] = bf16_rn(tanh_fast(z));
        }
        float zp[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < 4; t++) {
            mlp_v4f acc = {b2[t], b2[t], b2[t], b2[t]};
#pragma unroll
            for (int ks = 0; ks < 4; ks++) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[ks], bw[t][ks], acc, 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 4; j++) zp[j] = __builtin_fmaf(w3[t], tanh_fast(acc[j]), zp[j]);
        }
        // sum over the 16 output columns of the row (lanes 16 r .. 16 r + 15)
#pragma unroll
        for (int d = 8; d > 0; d >>= 1)
#pragma unroll
            for (int j = 0; j < 4; j++) zp[j] += __shfl_xor(zp[j], d);
        // lane c < 4 of row r writes sample 4 r + c of the tile
        const int o = 16 * tile + 4 * r + c;
        float zo = zp[0];
#pragma unroll
        for (int j = 1; j < 4; j++) zo = (c == j) ? zp[j] : zo;
        if (c < 4 && o < n) lam[o] = softplus_f(zo + b3);
    }
}

// ------------------------------------------------------------------ host

int Forecaster::ensure(int nsamp) {
    GPLUM_ENSURE(theta, MLP_NPARAM);
    GPLUM_ENSURE(grad, MLP_NPARAM);
    GPLUM_ENSURE(m, MLP_NPARAM);
    GPLUM_ENSURE(v, MLP_NPARAM);
    GPLUM_ENSURE(part, (size_t)MLP_NPARAM * ceil_div(nsamp, MLP_CHUNK));
    GPLUM_ENSURE(H1, (size_t)MLP_NH * nsamp);
    GPLUM_ENSURE(H2, (size_t)MLP_NH * nsamp);
    GPLUM_ENSURE(D1, (size_t)MLP_NH * nsamp);
    GPLUM_ENSURE(D2, (size_t)MLP_NH * nsamp);
    GPLUM_ENSURE(DZ, nsamp);
    GPLUM_ENSURE(loss_part, ceil_div(nsamp, 256) + 1);
    return GPLUM_OK;
}

int Forecaster::init_params(hipStream_t st, unsigned seed) {
    if (initialized) return GPLUM_OK;
    int rc = ensure(1);
    if (rc) return rc;
    std::vector<float> h(MLP_NPARAM, 0.f);
    // deterministic LCG + Box-Muller: W ~ N(0, 1/fan_in), biases 0, b3 = -2
    unsigned long long s = 0x9E3779B97F4A7C15ull ^ seed;
    auto uni = [&]() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return ((s >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    auto gauss = [&]() { return std::sqrt(-2.0 * std::log(uni())) * std::cos(6.283185307179586 * uni()); };
    for (int k = OFF_W1; k < OFF_B1; k++) h[k] = (float)(gauss() / 2.0);
    for (int k = OFF_W2; k < OFF_B2; k++) h[k] = (float)(gauss() / 8.0);
    for (int k = OFF_W3; k < OFF_B3; k++) h[k] = (float)(gauss() / 8.0);
    h[OFF_B3] = -2.f;
    GPLUM_HIP(hipMemcpyAsync(theta.p, h.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, st));
    GPLUM_HIP(hipMemsetAsync(m.p, 0, sizeof(float) * MLP_NPARAM, st));
    GPLUM_HIP(hipMemsetAsync(v.p, 0, sizeof(float) * MLP_NPARAM, st));
    adam_t = 0;
    initialized = true;
    return GPLUM_OK;
}

int Forecaster::forward(const float *d_x, int n, float *d_lam, hipStream_t st) {
    if (n <= 0) return GPLUM_OK;
    hipLaunchKernelGGL(k_mlp_forward, dim3(ceil_div(n, 256)), dim3(256), 0, st, theta.p, d_x, n, d_lam);
    GPLUM_HIP(hipGetLastError());
    return GPLUM_OK;
}

int Forecaster::predict(const float *d_x, int n, float *d_lam, hipStream_t st) {
    if (n <= 0) return GPLUM_OK;
    const int tiles = ceil_div(n, 16);
    const int blocks = std::min(ceil_div(tiles, MLP_PRED_BLOCK / 64), 2048);
    hipLaunchKernelGGL(k_mlp_predict, dim3(blocks), dim3(MLP_PRED_BLOCK), 0, st, theta.p, d_x, n, d_lam);
    GPLUM_HIP(hipGetLastError());
    return GPLUM_OK;
}

int Forecaster::loss_grad(const float *d_x, const float *d_y, int n, double *loss, hipStream_t st) {
    int rc = ensure(n);
    if (rc) return rc;
    const int nb = ceil_div(n, 256);
    hipLaunchKernelGGL(k_mlp_backward, dim3(nb), dim3(256), 0, st, theta.p, d_x, d_y, n, H1.p, H2.p, D1.p, D2.p,
                       DZ.p, loss_part.p);
    const int nchunk = ceil_div(n, MLP_CHUNK);
    hipLaunchKernelGGL(k_mlp_grad_part, dim3(ceil_div(MLP_NPARAM, 64), nchunk), dim3(64), 0, st, d_x, n, H1.p,
                       H2.p, D1.p, D2.p, DZ.p, part.p);
    hipLaunchKernelGGL(k_mlp_grad_sum, dim3(ceil_div(MLP_NPARAM, 256)), dim3(256), 0, st, nchunk, part.p, grad.p);
    GPLUM_HIP(hipGetLastError());
    if (loss) return read_loss(n, loss, st);
    return GPLUM_OK;
}

int Forecaster::read_loss(int n, double *loss, hipStream_t st) {
    const int nb = ceil_div(n, 256);
    std::vector<double> h(nb);
    GPLUM_HIP(hipMemcpyAsync(h.data(), loss_part.p, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    GPLUM_HIP(hipStreamSynchronize(st));
    double s = 0.;
    for (double x : h) s += x;
    *loss = s / n;
    return GPLUM_OK;
}

int Forecaster::adam_step(hipStream_t st) {
    adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)adam_t), c2 = 1.f - std::pow(b2, (float)adam_t);
    hipLaunchKernelGGL(k_mlp_adam, dim3(ceil_div(MLP_NPARAM, 256)), dim3(256), 0, st, theta.p, grad.p, m.p, v.p, lr,
                       b1, b2, c1, c2);
    GPLUM_HIP(hipGetLastError());
    return GPLUM_OK;
}

}  // namespace gplum

// ------------------------------------------------------------------ C ABI (tests)

namespace {
struct MlpCtx {
    hipStream_t st = nullptr;
    gplum::Forecaster f;
    gplum::DArray<float> x, y, lam;
};
thread_local MlpCtx t_mlp;

int mlp_ctx_init(MlpCtx &c) {
    if (!c.st) GPLUM_HIP(hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking));
    return GPLUM_OK;
}
}  // namespace

extern "C" int gplum_mlp_forward(const float *theta, const float *x, int n, float *lam) {
    if (!theta || (n > 0 && (!x || !lam)) || n < 0) return gplum::fail(GPLUM_E_ARG, "bad arguments");
    MlpCtx &c = t_mlp;
    int rc = mlp_ctx_init(c);
    if (rc) return rc;
    if ((rc = c.f.ensure(1))) return rc;
    GPLUM_ENSURE(c.x, 4 * (size_t)n + 4);
    GPLUM_ENSURE(c.lam, (size_t)n + 1);
    GPLUM_HIP(hipMemcpyAsync(c.f.theta.p, theta, sizeof(float) * gplum::MLP_NPARAM, hipMemcpyHostToDevice, c.st));
    GPLUM_HIP(hipMemcpyAsync(c.x.p, x, sizeof(float) * 4 * n, hipMemcpyHostToDevice, c.st));
    if ((rc = c.f.forward(c.x.p, n, c.lam.p, c.st))) return rc;
    GPLUM_HIP(hipMemcpyAsync(lam, c.lam.p, sizeof(float) * n, hipMemcpyDeviceToHost, c.st));
    GPLUM_HIP(hipStreamSynchronize(c.st));
    return GPLUM_OK;
}

extern "C" int gplum_mlp_predict(const float *theta, const float *x, int n, float *lam) {
    if (!theta || (n > 0 && (!x || !lam)) || n < 0) return gplum::fail(GPLUM_E_ARG, "bad arguments");
    MlpCtx &c = t_mlp;
    int rc = mlp_ctx_init(c);
    if (rc) return rc;
    if ((rc = c.f.ensure(1))) return rc;
    GPLUM_ENSURE(c.x, 4 * (size_t)n + 4);
    GPLUM_ENSURE(c.lam, (size_t)n + 1);
    GPLUM_HIP(hipMemcpyAsync(c.f.theta.p, theta, sizeof(float) * gplum::MLP_NPARAM, hipMemcpyHostToDevice, c.st));
    GPLUM_HIP(hipMemcpyAsync(c.x.p, x, sizeof(float) * 4 * n, hipMemcpyHostToDevice, c.st));
    if ((rc = c.f.predict(c.x.p, n, c.lam.p, c.st))) return rc;
    GPLUM_HIP(hipMemcpyAsync(lam, c.lam.p, sizeof(float) * n, hipMemcpyDeviceToHost, c.st));
    GPLUM_HIP(hipStreamSynchronize(c.st));
    return GPLUM_OK;
}

extern "C" int gplum_mlp_loss_grad(const float *theta, const float *x, const float *y, int n, double *loss,
                                   float *grad) {
    if (!theta || !x || !y || !grad || n <= 0) return gplum::fail(GPLUM_E_ARG, "bad arguments");
    MlpCtx &c = t_mlp;
    int rc = mlp_ctx_init(c);
    if (rc) return rc;
    if ((rc = c.f.ensure(n))) return rc;
    GPLUM_ENSURE(c.x, 4 * (size_t)n);
    GPLUM_ENSURE(c.y, (size_t)n);
    GPLUM_HIP(hipMemcpyAsync(c.f.theta.p, theta, sizeof(float) * gplum::MLP_NPARAM, hipMemcpyHostToDevice, c.st));
    GPLUM_HIP(hipMemcpyAsync(c.x.p, x, sizeof(float) * 4 * n, hipMemcpyHostToDevice, c.st));
    GPLUM_HIP(hipMemcpyAsync(c.y.p, y, sizeof(float) * n, hipMemcpyHostToDevice, c.st));
    if ((rc = c.f.loss_grad(c.x.p, c.y.p, n, loss, c.st))) return rc;
    GPLUM_HIP(hipMemcpyAsync(grad, c.f.grad.p, sizeof(float) * gplum::MLP_NPARAM, hipMemcpyDeviceToHost, c.st));
    GPLUM_HIP(hipStreamSynchronize(c.st));
    return GPLUM_OK;
}

extern "C" int gplum_mlp_train(float *theta, const float *x, const float *y, int n, int n_steps, float lr,
                               double *loss_first, double *loss_last) {
    if (!theta || !x || !y || n <= 0 || n_steps < 0 || !(lr > 0.f))
        return gplum::fail(GPLUM_E_ARG, "bad arguments");
    MlpCtx &c = t_mlp;
    int rc = mlp_ctx_init(c);
    if (rc) return rc;
    gplum::Forecaster f;   // fresh Adam state for this call
    if ((rc = f.ensure(n))) return rc;
    f.lr = lr;
    GPLUM_ENSURE(c.x, 4 * (size_t)n);
    GPLUM_ENSURE(c.y, (size_t)n);
    GPLUM_HIP(hipMemcpyAsync(f.theta.p, theta, sizeof(float) * gplum::MLP_NPARAM, hipMemcpyHostToDevice, c.st));
    GPLUM_HIP(hipMemsetAsync(f.m.p, 0, sizeof(float) * gplum::MLP_NPARAM, c.st));
    GPLUM_HIP(hipMemsetAsync(f.v.p, 0, sizeof(float) * gplum::MLP_NPARAM, c.st));
    GPLUM_HIP(hipMemcpyAsync(c.x.p, x, sizeof(float) * 4 * n, hipMemcpyHostToDevice, c.st));
    GPLUM_HIP(hipMemcpyAsync(c.y.p, y, sizeof(float) * n, hipMemcpyHostToDevice, c.st));
    for (int s = 0; s < n_steps; s++) {
        double *lp = (s == 0) ? loss_first : nullptr;
        if ((rc = f.loss_grad(c.x.p, c.y.p, n, lp, c.st))) return rc;
        if ((rc = f.adam_step(c.st))) return rc;
    }
    if (loss_last) {
        if ((rc = f.loss_grad(c.x.p, c.y.p, n, loss_last, c.st))) return rc;
    }
    GPLUM_HIP(hipMemcpyAsync(theta, f.theta.p, sizeof(float) * gplum::MLP_NPARAM, hipMemcpyDeviceToHost, c.st));
    GPLUM_HIP(hipStreamSynchronize(c.st));
    return GPLUM_OK;
}
