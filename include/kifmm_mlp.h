/*
 * kifmm_mlp.h -- the cost model's network on the GPU (C ABI).
 *
 * No counterpart in the reference (its nearest knobs are the hand-set
 * BS / UPBS block sizes, fmm/util.c:118-164).  A 16 -> 64 -> 64 -> 4
 * multilayer perceptron in FP32 (ReLU hidden layers, linear output) maps
 * 16 octree statistics of an input to the log phase times (up, U, V, down)
 * of one evaluation; the tuner uses it to pick points per box and the
 * precision path.  Forward, backward and Adam run as HIP kernels; every
 * reduction has a fixed order (deterministic).
 *
 * Parameter layout (KIFMM_MLP_PARAMS floats): W1 [16 x 64], b1 [64],
 * W2 [64 x 64], b2 [64], W3 [64 x 4], b3 [4], row-major (W[i][o] maps
 * input i to output o).  Batches are row-major [n x 16] inputs and
 * [n x 4] targets.  Functions return 0 or a negative KIFMM_E* code.
 */
#ifndef KIFMM_MLP_H
#define KIFMM_MLP_H

#include "kifmm_host.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { KIFMM_MLP_IN = 16, KIFMM_MLP_HIDDEN = 64, KIFMM_MLP_OUT = 4 };
enum { KIFMM_MLP_PARAMS = 16 * 64 + 64 + 64 * 64 + 64 + 64 * 4 + 4 };

typedef struct kifmm_mlp kifmm_mlp;

int kifmm_mlp_create(int device, const float* params, kifmm_mlp** out);
int kifmm_mlp_destroy(kifmm_mlp* h);
int kifmm_mlp_get_params(kifmm_mlp* h, float* out);
/* Y [n x 4] = network(X [n x 16]) */
int kifmm_mlp_predict(kifmm_mlp* h, const float* X, int n, float* Y);
/* gradient of mean((network(X) - T)^2) with respect to the parameters,
 * and that loss */
int kifmm_mlp_gradient(kifmm_mlp* h, const float* X, const float* T, int n, float* grad, double* loss);
/* epochs of full-batch Adam (beta 0.9 / 0.999, eps 1e-8) at rate lr;
 * losses[e] (may be NULL) is the loss before step e */
int kifmm_mlp_train(kifmm_mlp* h, const float* X, const float* T, int n, int epochs, float lr, double* losses);

#ifdef __cplusplus
}
#endif
#endif
