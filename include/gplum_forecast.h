/* gplum_forecast.h -- the neighbour-count forecaster (learned buffer sizing).
 *
 * No counterpart in the reference: GPLUM sizes its neighbour lists with std::vector
 * growth on the CPU (neighbor.h, hard.h).  On the GPU the FP32 soft-force kernel
 * writes neighbour candidates into a buffer that must be sized before the launch;
 * this small network predicts per-particle candidate counts so that the buffer is
 * sized ahead of time, while the exact search stays the judge (an undersized
 * buffer is detected and redone, so results never depend on the forecast).
 *
 * Network: 4 inputs -> 64 tanh -> 64 tanh -> 1, rate = softplus(output), trained
 * with the mean Poisson negative log-likelihood  mean(rate - y log(rate + 1e-6)).
 * Parameters theta[4545] (float32, row-major): W1[64][4], b1[64], W2[64][64],
 * b2[64], w3[64], b3.  Feature arrays are feature-major: x[q*n + i].
 */
#ifndef GPLUM_FORECAST_H
#define GPLUM_FORECAST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPLUM_MLP_NPARAM 4545

/* rate[i] = softplus(MLP(x[:, i])) for n samples.  Host pointers; synchronous. */
int gplum_mlp_forward(const float *theta, const float *x, int n, float *rate);

/* The rates of gplum_mlp_forward with the 64x64 hidden layer on the matrix cores
 * (MFMA, bf16 operands, FP32 accumulation) and hardware tanh: the forecast the
 * buffer sizing uses (within about 1e-2 relative of gplum_mlp_forward).  Host
 * pointers; synchronous. */
int gplum_mlp_predict(const float *theta, const float *x, int n, float *rate);

/* Mean Poisson NLL over n samples (targets y) and its gradient with respect to theta
 * (grad[4545], float32; summed over samples in FP64, deterministic).  loss may be
 * null.  Host pointers; synchronous. */
int gplum_mlp_loss_grad(const float *theta, const float *x, const float *y, int n, double *loss,
                        float *grad);

/* n_steps full-batch Adam steps (beta 0.9/0.999, eps 1e-8, learning rate lr) on the
 * device, updating theta in place.  loss_first / loss_last (may be null) receive the
 * loss before the first and after the last step.  Host pointers; synchronous. */
int gplum_mlp_train(float *theta, const float *x, const float *y, int n, int n_steps, float lr,
                    double *loss_first, double *loss_last);

#ifdef __cplusplus
}
#endif

#endif /* GPLUM_FORECAST_H */
