/*
 * dpd_mlp.h -- C ABI of the surrogate network of libospreydpd_amd.so.
 *
 * No counterpart in the reference, which has no learned component.  A small MLP
 * (nin -> H -> H -> nout, tanh hidden units) maps run parameters (a_ij, composition,
 * bond stiffness) to equilibrium observables (pressure, area per lipid).  Its matrix
 * products run on MI355X matrix cores in BF16 with FP32 accumulation
 * (v_mfma_f32_16x16x16_bf16); master weights, biases and the Adam state are FP32.
 * Inputs and targets are standardised with the statistics given to
 * dpd_mlp_set_scaling.  The simulation stays the check on every prediction.
 */
#ifndef DPD_MLP_H
#define DPD_MLP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct dpd_mlp dpd_mlp;

/* nin <= 16, nout <= 16, hidden a multiple of 16 (<= 1024).  Weights are drawn with
 * Glorot-uniform scaling from `seed`. */
int  dpd_mlp_create(int32_t nin, int32_t hidden, int32_t nout, uint64_t seed, int32_t device, dpd_mlp** out);
void dpd_mlp_destroy(dpd_mlp* m);
const char* dpd_mlp_last_error(const dpd_mlp* m);

/* x_std = (x - in_mean) / in_scale, y = out_mean + out_scale * net(x_std) */
int dpd_mlp_set_scaling(dpd_mlp* m, const double* in_mean, const double* in_scale, const double* out_mean,
                        const double* out_scale);

/* parameter count and FP32 access in the order W1[nin][H], b1[H], W2[H][H], b2[H],
 * W3[H][nout], b3[nout] */
int64_t dpd_mlp_param_count(const dpd_mlp* m);
int dpd_mlp_get_params(dpd_mlp* m, float* params);
int dpd_mlp_set_params(dpd_mlp* m, const float* params);

/* One forward + backward pass on n samples (raw units): loss = mean over samples and
 * outputs of 1/2 (net(x_std) - y_std)^2; grads in the parameter order above. */
int dpd_mlp_gradients(dpd_mlp* m, int64_t n, const double* x, const double* y, double* loss, float* grads);

/* Full-batch Adam (beta 0.9/0.999, eps 1e-8) for `iters` iterations at learning rate
 * lr; loss_out[iters] (may be NULL) receives the loss before each update. */
int dpd_mlp_train(dpd_mlp* m, int64_t n, const double* x, const double* y, int32_t iters, double lr,
                  double* loss_out);

/* predictions in raw units, y[n][nout] */
int dpd_mlp_predict(dpd_mlp* m, int64_t n, const double* x, double* y);

#ifdef __cplusplus
}
#endif
#endif
