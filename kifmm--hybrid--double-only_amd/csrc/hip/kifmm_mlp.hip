/*
 * kifmm_mlp.hip -- the cost model's network on the GPU: a 16 -> 64 -> 64 -> 4
 * multilayer perceptron in FP32 (ReLU hidden layers, linear output) with
 * forward, backward and Adam kernels.  It predicts the log phase times of
 * one evaluation from 16 octree statistics (computed on the host by
 * kifmm_problem_features); the tuner uses it to choose points per box and
 * the precision path.  No counterpart in the reference, whose nearest knobs
 * are the hand-set BS / UPBS block sizes (fmm/util.c:118-164).
 *
 * Parameters live in one device array in the order W1 [16 x 64], b1 [64],
 * W2 [64 x 64], b2 [64], W3 [64 x 4], b3 [4] (row-major, W[i][o] maps input
 * i to output o).  Every reduction runs in a fixed order, so forward,
 * gradients and training are deterministic.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "kifmm_mlp.h"
#include "../host/kifmm_internal.h"

using kifmm::set_error;

namespace {

constexpr int D0 = KIFMM_MLP_IN, D1 = KIFMM_MLP_HIDDEN, D2 = KIFMM_MLP_HIDDEN, D3 = KIFMM_MLP_OUT;
constexpr int OW1 = 0, OB1 = OW1 + D0 * D1, OW2 = OB1 + D1, OB2 = OW2 + D1 * D2, OW3 = OB2 + D2,
              OB3 = OW3 + D2 * D3, NPAR = OB3 + D3;
static_assert(NPAR == KIFMM_MLP_PARAMS, "parameter count");

/* Dense layer over a batch: out[n][o] = act(b[o] + sum_i in[n][i] W[i][o]);
 * pre-activations kept for the backward pass.  One thread per (n, o). */
__global__ void k_dense(int n, int din, int dout, const float* __restrict__ in, const float* __restrict__ W,
                        const float* __restrict__ b, float* __restrict__ pre, float* __restrict__ out, int relu) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * dout) return;
  const int r = idx / dout, o = idx % dout;
  float s = b[o];
  for (int i = 0; i < din; i++) s = __builtin_fmaf(in[(size_t)r * din + i], W[(size_t)i * dout + o], s);
  if (pre) pre[idx] = s;
  out[idx] = relu ? fmaxf(s, 0.0f) : s;
}

/* Mean-squared loss gradient: g = 2 (y - t) / (n * dout); loss partials per
 * sample row (summed on the host in order). */
__global__ void k_mse_grad(int n, int dout, const float* __restrict__ y, const float* __restrict__ t,
                           float* __restrict__ g, float* __restrict__ row_loss) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  float l = 0.0f;
  for (int o = 0; o < dout; o++) {
    const float d = y[(size_t)r * dout + o] - t[(size_t)r * dout + o];
    g[(size_t)r * dout + o] = 2.0f * d / (float)(n * dout);
    l = __builtin_fmaf(d, d, l);
  }
  row_loss[r] = l;
}

/* Weight and bias gradients of a dense layer: dW[i][o] = sum_n in[n][i] g[n][o],
 * db[o] = sum_n g[n][o]; one thread per parameter, samples in order. */
__global__ void k_dense_wgrad(int n, int din, int dout, const float* __restrict__ in, const float* __restrict__ g,
                              float* __restrict__ dW, float* __restrict__ db) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < din * dout) {
    const int i = idx / dout, o = idx % dout;
    float s = 0.0f;
    for (int r = 0; r < n; r++) s = __builtin_fmaf(in[(size_t)r * din + i], g[(size_t)r * dout + o], s);
    dW[idx] = s;
  } else if (idx < din * dout + dout) {
    const int o = idx - din * dout;
    float s = 0.0f;
    for (int r = 0; r < n; r++) s += g[(size_t)r * dout + o];
    db[o] = s;
  }
}

/* Gradient through a dense layer and the previous ReLU:
 * gin[n][i] = (pre_in[n][i] > 0) * sum_o g[n][o] W[i][o]. */
__global__ void k_dense_xgrad(int n, int din, int dout, const float* __restrict__ g, const float* __restrict__ W,
                              const float* __restrict__ pre_in, float* __restrict__ gin) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * din) return;
  const int r = idx / din, i = idx % din;
  float s = 0.0f;
  for (int o = 0; o < dout; o++) s = __builtin_fmaf(g[(size_t)r * dout + o], W[(size_t)i * dout + o], s);
  gin[idx] = pre_in[idx] > 0.0f ? s : 0.0f;
}

/* Adam (Kingma & Ba 2015) with bias correction, one thread per parameter. */
__global__ void k_adam(int np, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

inline int blocks(long n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

struct kifmm_mlp {
  int dev = 0;
  hipStream_t stream = nullptr;
  int cap = 0; /* batch capacity of the work buffers */
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  float *x = nullptr, *t = nullptr, *a1 = nullptr, *h1 = nullptr, *a2 = nullptr, *h2 = nullptr, *y = nullptr;
  float *gy = nullptr, *g2 = nullptr, *g1 = nullptr, *rl = nullptr;
  long step = 0;
};

namespace {

#define MLP_TRY(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      set_error(std::string("kifmm_mlp: ") + #expr + ": " + hipGetErrorString(e_)); \
      return KIFMM_EDEVICE;                                                    \
    }                                                                          \
  } while (0)

void release(kifmm_mlp* h) {
  float** bufs[] = {&h->x, &h->t, &h->a1, &h->h1, &h->a2, &h->h2, &h->y, &h->gy, &h->g2, &h->g1, &h->rl};
  for (float** b : bufs)
    if (*b) {
      (void)hipFree(*b);
      *b = nullptr;
    }
  h->cap = 0;
}

int reserve(kifmm_mlp* h, int n) {
  if (n <= h->cap) return KIFMM_OK;
  release(h);
  const int c = n;
  struct {
    float** p;
    long n;
  } want[] = {{&h->x, (long)c * D0}, {&h->t, (long)c * D3}, {&h->a1, (long)c * D1}, {&h->h1, (long)c * D1},
              {&h->a2, (long)c * D2}, {&h->h2, (long)c * D2}, {&h->y, (long)c * D3}, {&h->gy, (long)c * D3},
              {&h->g2, (long)c * D2}, {&h->g1, (long)c * D1}, {&h->rl, (long)c}};
  for (auto& w : want)
    if (hipMalloc((void**)w.p, sizeof(float) * w.n) != hipSuccess) {
      release(h);
      set_error("kifmm_mlp: device allocation failed");
      return KIFMM_ENOMEM;
    }
  h->cap = c;
  return KIFMM_OK;
}

int forward(kifmm_mlp* h, int n) {
  const float* P = h->p;
  hipLaunchKernelGGL(k_dense, dim3(blocks((long)n * D1, 256)), dim3(256), 0, h->stream, n, D0, D1, h->x, P + OW1,
                     P + OB1, h->a1, h->h1, 1);
  hipLaunchKernelGGL(k_dense, dim3(blocks((long)n * D2, 256)), dim3(256), 0, h->stream, n, D1, D2, h->h1, P + OW2,
                     P + OB2, h->a2, h->h2, 1);
  hipLaunchKernelGGL(k_dense, dim3(blocks((long)n * D3, 256)), dim3(256), 0, h->stream, n, D2, D3, h->h2, P + OW3,
                     P + OB3, (float*)nullptr, h->y, 0);
  MLP_TRY(hipGetLastError());
  return KIFMM_OK;
}

/* loss and gradient of the mean-squared error over n rows of h->x / h->t */
int backward(kifmm_mlp* h, int n, double* loss) {
  const float* P = h->p;
  float* G = h->g;
  hipLaunchKernelGGL(k_mse_grad, dim3(blocks(n, 256)), dim3(256), 0, h->stream, n, D3, h->y, h->t, h->gy, h->rl);
  hipLaunchKernelGGL(k_dense_wgrad, dim3(blocks((long)D2 * D3 + D3, 256)), dim3(256), 0, h->stream, n, D2, D3, h->h2,
                     h->gy, G + OW3, G + OB3);
  hipLaunchKernelGGL(k_dense_xgrad, dim3(blocks((long)n * D2, 256)), dim3(256), 0, h->stream, n, D2, D3, h->gy,
                     P + OW3, h->a2, h->g2);
  hipLaunchKernelGGL(k_dense_wgrad, dim3(blocks((long)D1 * D2 + D2, 256)), dim3(256), 0, h->stream, n, D1, D2, h->h1,
                     h->g2, G + OW2, G + OB2);
  hipLaunchKernelGGL(k_dense_xgrad, dim3(blocks((long)n * D1, 256)), dim3(256), 0, h->stream, n, D1, D2, h->g2,
                     P + OW2, h->a1, h->g1);
  hipLaunchKernelGGL(k_dense_wgrad, dim3(blocks((long)D0 * D1 + D1, 256)), dim3(256), 0, h->stream, n, D0, D1, h->x,
                     h->g1, G + OW1, G + OB1);
  MLP_TRY(hipGetLastError());
  if (loss) {
    std::vector<float> rl(n);
    MLP_TRY(hipMemcpyAsync(rl.data(), h->rl, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream));
    MLP_TRY(hipStreamSynchronize(h->stream));
    double s = 0.0;
    for (int r = 0; r < n; r++) s += rl[r];
    *loss = s / ((double)n * D3);
  }
  return KIFMM_OK;
}

int upload(kifmm_mlp* h, const float* X, const float* T, int n) {
  if (n <= 0 || !X) {
    set_error("kifmm_mlp: empty batch");
    return KIFMM_EINVAL;
  }
  int r = reserve(h, n);
  if (r != KIFMM_OK) return r;
  MLP_TRY(hipMemcpyAsync(h->x, X, sizeof(float) * n * D0, hipMemcpyHostToDevice, h->stream));
  if (T) MLP_TRY(hipMemcpyAsync(h->t, T, sizeof(float) * n * D3, hipMemcpyHostToDevice, h->stream));
  return KIFMM_OK;
}

}  // namespace

extern "C" {

int kifmm_mlp_create(int device, const float* params, kifmm_mlp** out) {
  if (!out || !params) {
    set_error("kifmm_mlp_create: null argument");
    return KIFMM_EINVAL;
  }
  *out = nullptr;
  kifmm_mlp* h = new (std::nothrow) kifmm_mlp();
  if (!h) return KIFMM_ENOMEM;
  h->dev = device;
  auto fail = [&](int code) {
    kifmm_mlp_destroy(h);
    return code;
  };
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&h->stream) != hipSuccess) {
    set_error("kifmm_mlp_create: device or stream");
    return fail(KIFMM_EDEVICE);
  }
  float** arrs[] = {&h->p, &h->g, &h->m, &h->v};
  for (float** a : arrs)
    if (hipMalloc((void**)a, sizeof(float) * NPAR) != hipSuccess || hipMemset(*a, 0, sizeof(float) * NPAR) != hipSuccess) {
      set_error("kifmm_mlp_create: device allocation failed");
      return fail(KIFMM_ENOMEM);
    }
  if (hipMemcpy(h->p, params, sizeof(float) * NPAR, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("kifmm_mlp_create: upload failed");
    return fail(KIFMM_EDEVICE);
  }
  *out = h;
  return KIFMM_OK;
}

int kifmm_mlp_destroy(kifmm_mlp* h) {
  if (!h) return KIFMM_OK;
  (void)hipSetDevice(h->dev);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  release(h);
  float** arrs[] = {&h->p, &h->g, &h->m, &h->v};
  for (float** a : arrs)
    if (*a) (void)hipFree(*a);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return KIFMM_OK;
}

int kifmm_mlp_get_params(kifmm_mlp* h, float* out) {
  if (!h || !out) return KIFMM_EINVAL;
  MLP_TRY(hipSetDevice(h->dev));
  MLP_TRY(hipMemcpyAsync(out, h->p, sizeof(float) * NPAR, hipMemcpyDeviceToHost, h->stream));
  MLP_TRY(hipStreamSynchronize(h->stream));
  return KIFMM_OK;
}

int kifmm_mlp_predict(kifmm_mlp* h, const float* X, int n, float* Y) {
  if (!h || !Y) return KIFMM_EINVAL;
  MLP_TRY(hipSetDevice(h->dev));
  int r = upload(h, X, nullptr, n);
  if (r == KIFMM_OK) r = forward(h, n);
  if (r != KIFMM_OK) return r;
  MLP_TRY(hipMemcpyAsync(Y, h->y, sizeof(float) * n * D3, hipMemcpyDeviceToHost, h->stream));
  MLP_TRY(hipStreamSynchronize(h->stream));
  return KIFMM_OK;
}

int kifmm_mlp_gradient(kifmm_mlp* h, const float* X, const float* T, int n, float* grad, double* loss) {
  if (!h || !T || !grad) return KIFMM_EINVAL;
  MLP_TRY(hipSetDevice(h->dev));
  int r = upload(h, X, T, n);
  if (r == KIFMM_OK) r = forward(h, n);
  if (r == KIFMM_OK) r = backward(h, n, loss);
  if (r != KIFMM_OK) return r;
  MLP_TRY(hipMemcpyAsync(grad, h->g, sizeof(float) * NPAR, hipMemcpyDeviceToHost, h->stream));
  MLP_TRY(hipStreamSynchronize(h->stream));
  return KIFMM_OK;
}

int kifmm_mlp_train(kifmm_mlp* h, const float* X, const float* T, int n, int epochs, float lr, double* losses) {
  if (!h || !T || epochs < 1 || !(lr > 0.0f)) {
    set_error("kifmm_mlp_train: invalid argument");
    return KIFMM_EINVAL;
  }
  MLP_TRY(hipSetDevice(h->dev));
  int r = upload(h, X, T, n);
  if (r != KIFMM_OK) return r;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int e = 0; e < epochs; e++) {
    r = forward(h, n);
    if (r == KIFMM_OK) r = backward(h, n, losses ? &losses[e] : nullptr);
    if (r != KIFMM_OK) return r;
    h->step++;
    const float c1 = 1.0f - std::pow(b1, (float)h->step), c2 = 1.0f - std::pow(b2, (float)h->step);
    hipLaunchKernelGGL(k_adam, dim3(blocks(NPAR, 256)), dim3(256), 0, h->stream, NPAR, h->p, h->g, h->m, h->v, lr, b1,
                       b2, eps, c1, c2);
    MLP_TRY(hipGetLastError());
  }
  MLP_TRY(hipStreamSynchronize(h->stream));
  return KIFMM_OK;
}

}  // extern "C"
