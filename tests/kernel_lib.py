"""ctypes signatures of the kernel-level test entry points (include/kifmm_kernels.h),
built into the test-only build/libkifmm_ktest.so (not the product library)."""
import ctypes as C

import numpy as np

import kifmm_amd as K

KTEST = K.HERE / "build" / "libkifmm_ktest.so"
if not KTEST.exists():
    raise ImportError(f"kernel test library not built: {KTEST} (run __graft_entry__.build())")
lib = C.CDLL(str(KTEST))
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
for name, args in (
    ("kifmm_test_mfma_layout", [_dp, _dp, _dp]),
    ("kifmm_test_mfma_layout_f32", [C.POINTER(C.c_float)] * 3),
    ("kifmm_test_mfma_layout_bf16", [C.POINTER(C.c_float)] * 3),
    ("kifmm_test_gemm", [C.c_int, C.c_int, _dp, C.c_int, _dp, _ip, _dp, C.c_int]),
    ("kifmm_test_rsqrt", [C.c_int, _dp, _dp, _dp]),
    ("kifmm_test_rsq_clamp_f32", [C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("kifmm_test_dft_fwd", [C.c_int, C.c_int, _dp, _ip, _dp]),
    ("kifmm_test_dft_inv", [C.c_int, C.c_int, _dp, _dp]),
    ("kifmm_test_translation_spectra", [C.c_int, _dp]),
):
    f = getattr(lib, name)
    f.restype = C.c_int
    f.argtypes = args


def dp(a):
    return a.ctypes.data_as(_dp)


def ip(a):
    return a.ctypes.data_as(_ip)


def mfma_layout(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.zeros(256)
    K._check(lib.kifmm_test_mfma_layout(dp(a), dp(b), dp(out)))
    return out.reshape(64, 4)


def mfma_layout_f32(a, b):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    out = np.zeros(256, dtype=np.float32)
    fp = C.POINTER(C.c_float)
    K._check(lib.kifmm_test_mfma_layout_f32(a.ctypes.data_as(fp), b.ctypes.data_as(fp), out.ctypes.data_as(fp)))
    return out.reshape(64, 4)


def mfma_layout_bf16(a, b):
    """a, b: 64 x 8 per-lane values; returns the 64 x 4 accumulator elements"""
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(512)
    b = np.ascontiguousarray(b, dtype=np.float32).reshape(512)
    out = np.zeros(256, dtype=np.float32)
    fp = C.POINTER(C.c_float)
    K._check(lib.kifmm_test_mfma_layout_bf16(a.ctypes.data_as(fp), b.ctypes.data_as(fp), out.ctypes.data_as(fp)))
    return out.reshape(64, 4)


def gemm(A, B, sexp, C0=None, accumulate=False):
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    M, Kd = A.shape
    ncols = B.shape[1]
    Cm = np.zeros((M, ncols), order="F") if C0 is None else np.asfortranarray(C0, dtype=np.float64).copy(order="F")
    se = np.ascontiguousarray(sexp, dtype=np.int32)
    K._check(lib.kifmm_test_gemm(M, Kd, dp(A), ncols, dp(B), ip(se), dp(Cm), int(accumulate)))
    return Cm


def rsqrt(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    ym = np.zeros_like(x)
    K._check(lib.kifmm_test_rsqrt(len(x), dp(x), dp(y), dp(ym)))
    return y, ym


def rsq_clamp_f32(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros_like(x)
    fp = lambda a: a.ctypes.data_as(K.C.POINTER(K.C.c_float))
    K._check(lib.kifmm_test_rsq_clamp_f32(len(x), fp(x), fp(y)))
    return y


def grid_f(np_):
    N = 2 * np_
    return N * N * (np_ + 1)


def dft_fwd(np_, den, depth):
    """den [nblocks*8, n_e] -> spectra [nblocks, F, 8] complex"""
    den = np.ascontiguousarray(den, dtype=np.float64)
    depth = np.ascontiguousarray(depth, dtype=np.int32)
    nb = den.shape[0] // 8
    out = np.zeros((nb, grid_f(np_), 8), dtype=np.complex128)
    K._check(lib.kifmm_test_dft_fwd(np_, nb, dp(den), ip(depth), out.ctypes.data_as(_dp)))
    return out


def dft_inv(np_, spec):
    spec = np.ascontiguousarray(spec, dtype=np.complex128)
    nb = spec.shape[0]
    out = np.zeros((nb * 8, K.n_e(np_)))
    K._check(lib.kifmm_test_dft_inv(np_, nb, spec.ctypes.data_as(_dp), dp(out)))
    return out


def translation_spectra(np_):
    out = np.zeros((grid_f(np_), 343), dtype=np.complex128)
    K._check(lib.kifmm_test_translation_spectra(np_, out.ctypes.data_as(_dp)))
    return out
