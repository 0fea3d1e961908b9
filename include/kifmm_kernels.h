/*
 * kifmm_kernels.h -- kernel-level entry points used by the test suite.
 *
 * No counterpart in the reference: each call runs one evaluator kernel on
 * the current HIP device with host-staged inputs, exactly as
 * kifmm_gpu_evaluate launches it, so that stage can be checked on its own
 * (against numpy or the CPU oracle).  They allocate and free device memory
 * per call and are not meant for production loops.  Return 0 or a negative
 * KIFMM_E* code.
 */
#ifndef KIFMM_KERNELS_H
#define KIFMM_KERNELS_H

#include "kifmm_host.h"

#ifdef __cplusplus
extern "C" {
#endif

/* v_mfma_f64_16x16x4f64 on one wave: lane l supplies a[l] and b[l];
 * out[4*l + r] is accumulator element r of lane l. */
int kifmm_test_mfma_layout(const double* a, const double* b, double* out);

/* v_mfma_f32_16x16x4f32 on one wave, same convention (FP32). */
int kifmm_test_mfma_layout_f32(const float* a, const float* b, float* out);
/* v_mfma_f32_16x16x32_bf16 probe: lane l passes a[8l..8l+7], b[8l..8l+7]
 * (converted to bf16); out[4l + r] = accumulator element r of lane l */
int kifmm_test_mfma_layout_bf16(const float* a, const float* b, float* out);

/* C[:, j] = (acc ? C[:, j] : 0) + 2^sexp[j] * A * B[:, j] through the
 * evaluator's FP64 MFMA GEMM; A [M x K], B [K x ncols], C [M x ncols],
 * all column-major with leading dimensions M, K, M. */
int kifmm_test_gemm(int M, int K, const double* A, int ncols, const double* B, const int* sexp,
                    double* C, int accumulate);

/* y = x^(-1/2) as the point kernels compute it; ymask is the r == 0-masked
 * form (0 where x == 0). */
int kifmm_test_rsqrt(int n, const double* x, double* y, double* ymask);
/* y = min(1, 1/sqrt|x|) in FP32 as the mixed U-list forms it (v_rsq_f32 with
 * the |x| and clamp modifiers) */
int kifmm_test_rsq_clamp_f32(int n, const float* x, float* y);

/* Forward transform of nblocks groups of 8 equivalent densities
 * (den[(8b+o)*n_e + j], scaled by 2^depth[8b+o] on the grid): spectra are
 * written as interleaved complex [b][f][o], f = (kz*N + ky)*(N/2+1) + kx. */
int kifmm_test_dft_fwd(int np, int nblocks, const double* den, const int* depth, double* spectra);

/* Inverse: spectra as above -> check-surface values [(8b+o)*n_e + j],
 * normalised by 1/N^3. */
int kifmm_test_dft_inv(int np, int nblocks, const double* spectra, double* values);

/* The 343 translation spectra, interleaved complex [f][id]. */
int kifmm_test_translation_spectra(int np, double* T);

#ifdef __cplusplus
}
#endif
#endif
