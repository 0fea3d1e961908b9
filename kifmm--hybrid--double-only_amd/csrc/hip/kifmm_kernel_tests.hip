/*
 * kifmm_kernel_tests.hip -- kernel-level entry points for the test suite.
 *
 * Each function stages host arrays on the current device, runs one of the
 * evaluator's kernels exactly as kifmm_gpu_evaluate launches it, copies the
 * result back and frees everything.  They have no counterpart in the
 * reference; they exist so each stage can be checked on its own against
 * numpy and the CPU oracle.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "kifmm_kernels.h"
#include "../host/kifmm_internal.h"
#include "kernels.hpp"

using namespace kifmm_dev;
using kifmm::set_error;

namespace {

struct DevBufs {
  std::vector<void*> p;
  ~DevBufs() {
    for (void* q : p) (void)hipFree(q);
  }
  template <typename T>
  T* up(const T* src, size_t n) {
    void* d = nullptr;
    if (hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
    p.push_back(d);
    if (src) {
      if (hipMemcpy(d, src, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    } else if (hipMemset(d, 0, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
      return nullptr;
    }
    return (T*)d;
  }
};

int finish(const char* what) {
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return KIFMM_EDEVICE;
  }
  return KIFMM_OK;
}

template <typename T>
int down(T* dst, const T* src, size_t n) {
  if (hipMemcpy(dst, src, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("device to host copy failed");
    return KIFMM_EDEVICE;
  }
  return KIFMM_OK;
}

/* one wave: lane l passes a[l], b[l] to v_mfma_f64_16x16x4f64 and stores
 * the 4 accumulator elements it receives */
__global__ void k_mfma_probe(const double* a, const double* b, double* out) {
  const int l = threadIdx.x;
  v4d acc = {0.0, 0.0, 0.0, 0.0};
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[l], b[l], acc, 0, 0, 0);
  for (int r = 0; r < 4; r++) out[l * 4 + r] = acc[r];
}

__global__ void k_rsqrt_probe(const double* x, double* y, double* ymask, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    y[i] = rsqrt_fp64(x[i]);
    ymask[i] = inv_r<true>(x[i]);
  }
}

/* the clamped reciprocal square root of the mixed U-list: min(1, 1/sqrt|x|) */
__global__ void k_rsq_clamp_probe(const float* x, float* y, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = __builtin_amdgcn_fmed3f(__builtin_amdgcn_rsqf(__builtin_fabsf(x[i])), 0.0f, 1.0f);
}

constexpr int rup(int x, int m) { return (x + m - 1) / m * m; }

template <int NP>
int dft_fwd_np(int nblocks, const double* den, const int* depth, double* out) {
  using G = Grid<NP>;
  const int ne = G::NE, lde = rup(ne, 8), N = G::N, F = G::F;
  DevBufs B;
  std::vector<double> u((size_t)(1 + 8 * nblocks) * lde, 0.0);
  std::vector<int> dep(1 + 8 * nblocks, 0), blocks(nblocks), slots(nblocks);
  for (int b = 0; b < nblocks; b++) {
    blocks[b] = 1 + 8 * b;
    slots[b] = b + 1;
    for (int o = 0; o < 8; o++) {
      dep[1 + 8 * b + o] = depth[8 * b + o];
      for (int j = 0; j < ne; j++) u[(size_t)(1 + 8 * b + o) * lde + j] = den[(size_t)(8 * b + o) * ne + j];
    }
  }
  std::vector<double2> tw(N);
  for (int m = 0; m < N; m++) {
    const long double a = 2.0L * 3.141592653589793238462643383279502884L * m / N;
    tw[m] = make_double2((double)cosl(a), (double)-sinl(a));
  }
  std::vector<int> c2s(NP * NP * NP, -1);
  int cnt = 0;
  for (int i = 0; i < NP; i++)
    for (int j = 0; j < NP; j++)
      for (int k = 0; k < NP; k++)
        if (i == 0 || i == NP - 1 || j == 0 || j == NP - 1 || k == 0 || k == NP - 1) c2s[(i * NP + j) * NP + k] = cnt++;
  int* dblk = B.up(blocks.data(), blocks.size());
  int* dslot = B.up(slots.data(), slots.size());
  int* ddep = B.up(dep.data(), dep.size());
  double* du = B.up(u.data(), u.size());
  double2* dtw = B.up(tw.data(), tw.size());
  int* dc2s = B.up(c2s.data(), c2s.size());
  double2* dX = B.up<double2>(nullptr, (size_t)(nblocks + 1) * F * 8);
  if (!dblk || !dslot || !ddep || !du || !dtw || !dc2s || !dX) {
    set_error("device allocation failed");
    return KIFMM_ENOMEM;
  }
  hipLaunchKernelGGL(k_dft_fwd<NP>, dim3(nblocks), dim3(512), 0, 0, dblk, dslot, ddep, du, lde, dtw, dc2s, dX);
  int r = finish("k_dft_fwd");
  if (r != KIFMM_OK) return r;
  return down((double2*)out, dX + (size_t)F * 8, (size_t)nblocks * F * 8);
}

template <int NP>
int dft_inv_np(int nblocks, const double* spec, double* out) {
  using G = Grid<NP>;
  const int ne = G::NE, lde = rup(ne, 8), N = G::N, F = G::F;
  DevBufs B;
  std::vector<int> blocks(nblocks), slots(nblocks);
  for (int b = 0; b < nblocks; b++) {
    blocks[b] = 1 + 8 * b;
    slots[b] = b + 1;
  }
  std::vector<double2> tw(N);
  for (int m = 0; m < N; m++) {
    const long double a = 2.0L * 3.141592653589793238462643383279502884L * m / N;
    tw[m] = make_double2((double)cosl(a), (double)-sinl(a));
  }
  std::vector<int> s2c;
  for (int i = 0; i < NP; i++)
    for (int j = 0; j < NP; j++)
      for (int k = 0; k < NP; k++)
        if (i == 0 || i == NP - 1 || j == 0 || j == NP - 1 || k == 0 || k == NP - 1) s2c.push_back((i * NP + j) * NP + k);
  std::vector<double2> Y((size_t)(nblocks + 1) * F * 8);
  for (size_t i = 0; i < (size_t)nblocks * F * 8; i++) Y[(size_t)F * 8 + i] = make_double2(spec[2 * i], spec[2 * i + 1]);
  int* dblk = B.up(blocks.data(), blocks.size());
  int* dslot = B.up(slots.data(), slots.size());
  double2* dtw = B.up(tw.data(), tw.size());
  int* ds2c = B.up(s2c.data(), s2c.size());
  double2* dY = B.up(Y.data(), Y.size());
  double* dchk = B.up<double>(nullptr, (size_t)(1 + 8 * nblocks) * lde);
  if (!dblk || !dslot || !dtw || !ds2c || !dY || !dchk) {
    set_error("device allocation failed");
    return KIFMM_ENOMEM;
  }
  hipLaunchKernelGGL(k_dft_inv<NP>, dim3(nblocks), dim3(512), 0, 0, dblk, dslot, dY, dtw, ds2c, dchk, lde);
  int r = finish("k_dft_inv");
  if (r != KIFMM_OK) return r;
  std::vector<double> h((size_t)(1 + 8 * nblocks) * lde);
  r = down(h.data(), dchk, h.size());
  if (r != KIFMM_OK) return r;
  for (int n = 0; n < 8 * nblocks; n++)
    for (int j = 0; j < ne; j++) out[(size_t)n * ne + j] = h[(size_t)(1 + n) * lde + j];
  return KIFMM_OK;
}

template <int NP>
int trans_np(double* out) {
  using G = Grid<NP>;
  const int N = G::N, F = G::F;
  DevBufs B;
  std::vector<double2> tw(N);
  for (int m = 0; m < N; m++) {
    const long double a = 2.0L * 3.141592653589793238462643383279502884L * m / N;
    tw[m] = make_double2((double)cosl(a), (double)-sinl(a));
  }
  double2* dtw = B.up(tw.data(), tw.size());
  double2* dT = B.up<double2>(nullptr, (size_t)F * 343);
  if (!dtw || !dT) {
    set_error("device allocation failed");
    return KIFMM_ENOMEM;
  }
  hipLaunchKernelGGL(k_trans<NP>, dim3(343), dim3(256), 0, 0, dtw, dT);
  int r = finish("k_trans");
  if (r != KIFMM_OK) return r;
  return down((double2*)out, dT, (size_t)F * 343);
}

}  // namespace

extern "C" {

namespace {
__global__ void k_mfma32_probe(const float* a, const float* b, float* out) {
  typedef float v4f __attribute__((ext_vector_type(4)));
  const int l = threadIdx.x;
  v4f c = {0.0f, 0.0f, 0.0f, 0.0f};
  c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[l], b[l], c, 0, 0, 0);
  for (int r = 0; r < 4; r++) out[4 * l + r] = c[r];
}
/* one wave: lane l converts a[8l .. 8l+7], b[8l .. 8l+7] to bf16, passes them
 * to v_mfma_f32_16x16x32_bf16 and stores the 4 accumulator elements */
__global__ void k_mfma_bf16_probe(const float* a, const float* b, float* out) {
  typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
  typedef float v4f __attribute__((ext_vector_type(4)));
  const int l = threadIdx.x;
  v8bf va, vb;
  for (int e = 0; e < 8; e++) {
    va[e] = (__bf16)a[8 * l + e];
    vb[e] = (__bf16)b[8 * l + e];
  }
  v4f c = {0.0f, 0.0f, 0.0f, 0.0f};
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va, vb, c, 0, 0, 0);
  for (int r = 0; r < 4; r++) out[4 * l + r] = c[r];
}
}  // namespace

int kifmm_test_mfma_layout_bf16(const float* a, const float* b, float* out) {
  if (!a || !b || !out) return KIFMM_EINVAL;
  DevBufs B;
  float* da = B.up(a, 512);
  float* db = B.up(b, 512);
  float* dout = B.up<float>(nullptr, 256);
  if (!da || !db || !dout) {
    set_error("device allocation failed");
    return KIFMM_ENOMEM;
  }
  hipLaunchKernelGGL(k_mfma_bf16_probe, dim3(1), dim3(64), 0, 0, da, db, dout);
  const int r = finish("k_mfma_bf16_probe");
  if (r != KIFMM_OK) return r;
  return down(out, dout, 256);
}

int kifmm_test_mfma_layout_f32(const float* a, const float* b, float* out) {
  if (!a || !b || !out) return KIFMM_EINVAL;
  DevBufs B;
  float* da = B.up(a, 64);
  float* db = B.up(b, 64);
  float* dout = B.up<float>(nullptr, 256);
  if (!da || !db || !dout) {
    set_error("device allocation failed");
    return KIFMM_ENOMEM;
  }
  hipLaunchKernelGGL(k_mfma32_probe, dim3(1), dim3(64), 0, 0, da, db, dout);
  int r = finish("k_mfma32_probe");
  if (r != KIFMM_OK) return r;
  return down(out, dout, 256);
}

int kifmm_test_mfma_layout(const double* a, const double* b, double* out) {
  if (!a || !b || !out) return KIFMM_EINVAL;
  DevBufs B;
  double* da = B.up(a, 64);
  double* db = B.up(b, 64);
  double* dout = B.up<double>(nullptr, 256);
  if (!da || !db || !dout) {
    set_error("device allocation failed");
    return KIFMM_ENOMEM;
  }
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, 0, da, db, dout);
  int r = finish("k_mfma_probe");
  if (r != KIFMM_OK) return r;
  return down(out, dout, 256);
}

int kifmm_test_gemm(int M, int K, const double* A, int ncols, const double* Bm, const int* sexp, double* Cm,
                    int accumulate) {
  if (M <= 0 || K <= 0 || ncols <= 0 || !A || !Bm || !sexp || !Cm) {
    set_error("kifmm_test_gemm: invalid argument");
    return KIFMM_EINVAL;
  }
  const int lda = rup(M, 16), Kp = rup(K, 4);
  std::vector<double> Ap((size_t)lda * Kp, 0.0), Bp((size_t)Kp * ncols, 0.0);
  for (int k = 0; k < K; k++)
    for (int i = 0; i < M; i++) Ap[(size_t)k * lda + i] = A[(size_t)k * M + i];
  for (int c = 0; c < ncols; c++)
    for (int k = 0; k < K; k++) Bp[(size_t)c * Kp + k] = Bm[(size_t)c * K + k];
  std::vector<int> in(ncols), out(ncols);
  for (int c = 0; c < ncols; c++) {
    in[c] = c * Kp;
    out[c] = c * M;
  }
  DevBufs B;
  double* dA = B.up(Ap.data(), Ap.size());
  double* dB = B.up(Bp.data(), Bp.size());
  double* dC = B.up(Cm, (size_t)M * ncols);
  int* din = B.up(in.data(), in.size());
  int* dout = B.up(out.data(), out.size());
  int* dse = B.up(sexp, ncols);
  if (!dA || !dB || !dC || !din || !dout || !dse) {
    set_error("device allocation failed");
    return KIFMM_ENOMEM;
  }
  /* the evaluator's shape rule and launcher */
  const GemmShape sh = gemm_shape(ncols, M, Kp);
  double* dws = sh.nsplit > 1 ? B.up<double>(nullptr, (size_t)sh.nsplit * ncols * sh.ldw) : nullptr;
  if (sh.nsplit > 1 && !dws) {
    set_error("device allocation failed");
    return KIFMM_ENOMEM;
  }
  if (launch_gemm_cols(0, sh, M, Kp, dA, lda, ncols, din, dout, dse, dB, dC, accumulate, dws) != hipSuccess) {
    set_error("GEMM launch failed");
    return KIFMM_EDEVICE;
  }
  int r = finish("k_gemm_cols");
  if (r != KIFMM_OK) return r;
  return down(Cm, dC, (size_t)M * ncols);
}

int kifmm_test_rsqrt(int n, const double* x, double* y, double* ymask) {
  if (n <= 0 || !x || !y || !ymask) return KIFMM_EINVAL;
  DevBufs B;
  double* dx = B.up(x, n);
  double* dy = B.up<double>(nullptr, n);
  double* dm = B.up<double>(nullptr, n);
  if (!dx || !dy || !dm) {
    set_error("device allocation failed");
    return KIFMM_ENOMEM;
  }
  hipLaunchKernelGGL(k_rsqrt_probe, dim3((n + 255) / 256), dim3(256), 0, 0, dx, dy, dm, n);
  int r = finish("k_rsqrt_probe");
  if (r != KIFMM_OK) return r;
  r = down(y, dy, n);
  if (r != KIFMM_OK) return r;
  return down(ymask, dm, n);
}

int kifmm_test_rsq_clamp_f32(int n, const float* x, float* y) {
  if (n <= 0 || !x || !y) return KIFMM_EINVAL;
  DevBufs B;
  float* dx = B.up(x, n);
  float* dy = B.up<float>(nullptr, n);
  if (!dx || !dy) {
    set_error("device allocation failed");
    return KIFMM_ENOMEM;
  }
  hipLaunchKernelGGL(k_rsq_clamp_probe, dim3((n + 255) / 256), dim3(256), 0, 0, dx, dy, n);
  const int r = finish("k_rsq_clamp_probe");
  if (r != KIFMM_OK) return r;
  return down(y, dy, n);
}

int kifmm_test_dft_fwd(int np, int nblocks, const double* den, const int* depth, double* spectra) {
  if (nblocks <= 0 || !den || !depth || !spectra) return KIFMM_EINVAL;
  switch (np) {
    case 3: return dft_fwd_np<3>(nblocks, den, depth, spectra);
    case 4: return dft_fwd_np<4>(nblocks, den, depth, spectra);
    case 6: return dft_fwd_np<6>(nblocks, den, depth, spectra);
    default: set_error("NP must be 3, 4 or 6"); return KIFMM_EUNSUP;
  }
}

int kifmm_test_dft_inv(int np, int nblocks, const double* spectra, double* values) {
  if (nblocks <= 0 || !spectra || !values) return KIFMM_EINVAL;
  switch (np) {
    case 3: return dft_inv_np<3>(nblocks, spectra, values);
    case 4: return dft_inv_np<4>(nblocks, spectra, values);
    case 6: return dft_inv_np<6>(nblocks, spectra, values);
    default: set_error("NP must be 3, 4 or 6"); return KIFMM_EUNSUP;
  }
}

int kifmm_test_translation_spectra(int np, double* T) {
  if (!T) return KIFMM_EINVAL;
  switch (np) {
    case 3: return trans_np<3>(T);
    case 4: return trans_np<4>(T);
    case 6: return trans_np<6>(T);
    default: set_error("NP must be 3, 4 or 6"); return KIFMM_EUNSUP;
  }
}

}  // extern "C"
