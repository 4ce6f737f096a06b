/*
 * helix_mlp.h -- C-ABI of the learned-closure network (libhelix_hip.so).
 *
 * No counterpart in the reference: the reference truncates the hierarchy plainly (the deepest tier has
 * no up-neighbours, src/liouville.cu:818-823 reads the zero sentinel block).  The network replaces the
 * missing tier-(L+1) contribution of the deepest kept tier.
 *
 * A 3-layer perceptron in -> hidden1 -> hidden2 -> out with tanh on both hidden layers, BF16 weights,
 * FP32 biases, FP32 activations and FP32 accumulation; FP32 master weights updated by Adam
 * (beta1 0.9, beta2 0.999, eps 1e-8).  Parameters are laid out [W1 b1 W2 b2 W3 b3], W stored
 * [out][in] row-major.  Loss: mean squared error over all outputs.  Return codes as helix_hip.h.
 */
#ifndef HELIX_MLP_H
#define HELIX_MLP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct helix_hip_mlp helix_hip_mlp;

int helix_hip_mlp_create(int in, int hidden1, int hidden2, int out, uint64_t seed, int device, helix_hip_mlp** mlp);
size_t helix_hip_mlp_param_count(const helix_hip_mlp* mlp);
/* sets the FP32 master parameters (and their BF16 copies); resets the optimiser state */
int helix_hip_mlp_set_params(helix_hip_mlp* mlp, const float* params);
int helix_hip_mlp_get_params(helix_hip_mlp* mlp, float* params);
/* y[n][out] = f(x[n][in]) */
int helix_hip_mlp_forward(helix_hip_mlp* mlp, const float* x, int n, float* y);
/* gradient of the MSE loss over the n samples with respect to every parameter (may be NULL) and the loss */
int helix_hip_mlp_gradient(helix_hip_mlp* mlp, const float* x, const float* target, int n, float* grad, float* loss);
/* minibatch Adam over the data set (kept on the device), `epochs` passes, rows shuffled per epoch by
 * SplitMix64(shuffle_seed); epoch_loss[e] (may be NULL) = mean loss of epoch e */
int helix_hip_mlp_train(helix_hip_mlp* mlp, const float* x, const float* target, int n, int epochs, int batch, float lr,
	uint64_t shuffle_seed, float* epoch_loss);
void helix_hip_mlp_destroy(helix_hip_mlp* mlp);
const char* helix_hip_mlp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
