// mlp.hpp -- device-side neighbour-count forecaster (4-64-64-1 MLP, FP32).
#pragma once
#include <cmath>

#include "engine.hpp"

namespace gplum {

constexpr int MLP_NIN = 4, MLP_NH = 64;
constexpr int MLP_NPARAM = MLP_NH * MLP_NIN + MLP_NH + MLP_NH * MLP_NH + MLP_NH + MLP_NH + 1;   // 4545

struct Forecaster {
    DArray<float> theta, grad, m, v;        // parameters, gradient, Adam moments
    DArray<float> H1, H2, D1, D2, DZ;       // per-sample activations / deltas, [n][64]
    DArray<double> loss_part, part;         // loss per block, gradient per sample chunk
    int adam_t = 0;
    float lr = 3e-3f;
    bool initialized = false;

    int ensure(int nsamp);
    int init_params(hipStream_t st, unsigned seed);
    // lam[i] = softplus(MLP(x[:, i])); x is feature-major [4][n]
    int forward(const float *d_x, int n, float *d_lam, hipStream_t st);
    // the same rates with the hidden layer on the matrix cores (bf16 operands, FP32
    // accumulation) and hardware tanh: for sizing, not for training
    int predict(const float *d_x, int n, float *d_lam, hipStream_t st);
    // fills grad with d(mean Poisson NLL)/d theta; loss (optional) synchronises the stream
    int loss_grad(const float *d_x, const float *d_y, int n, double *loss, hipStream_t st);
    // the mean loss of the last loss_grad call over its n samples (synchronises st)
    int read_loss(int n, double *loss, hipStream_t st);
    int adam_step(hipStream_t st);
};

}  // namespace gplum
