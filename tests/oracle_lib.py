"""ctypes access to the CPU oracle (oracle/liboracle.so) -- test infrastructure only.

The oracle restates the reference CPU path; it is the checker for the GPU
path and never part of the product.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "oracle" / "liboracle.so"
KERNEL_EXACT, KERNEL_REF = 0, 1

if not LIB.exists():
    raise ImportError(f"oracle not built: {LIB} (run __graft_entry__.build())")
lib = C.CDLL(str(LIB))
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


_sig("oracle_create", C.c_void_p, [C.c_int, C.c_char_p, C.c_int, C.c_int, C.c_long, C.c_int])
_sig("oracle_create_from_points", C.c_void_p, [C.c_int] + [_dp] * 7 + [C.c_int, C.c_int])
_sig("oracle_destroy", None, [C.c_void_p])
_sig("oracle_num_nodes", C.c_int, [C.c_void_p])
_sig("oracle_num_levels", C.c_int, [C.c_void_p])
_sig("oracle_level_starts", C.c_int, [C.c_void_p, _ip])
_sig("oracle_tree_arrays", C.c_int, [C.c_void_p] + [_ip] * 9)
_sig("oracle_list_csr", C.c_int, [C.c_void_p, C.c_int, _ip, _ip])
_sig("oracle_list_nnz", C.c_int, [C.c_void_p, C.c_int])
_sig("oracle_points", C.c_int, [C.c_void_p] + [_dp] * 7)
_sig("oracle_perms", C.c_int, [C.c_void_p, _ip, _ip])
_sig("oracle_build_operators", C.c_int, [C.c_void_p, C.c_int])
_sig("oracle_set_operators", C.c_int, [C.c_void_p] + [_dp] * 4)
_sig("oracle_get_operators", C.c_int, [C.c_void_p] + [_dp] * 4)
_sig("oracle_evaluate", C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _dp])
_sig("oracle_intermediates", C.c_int, [C.c_void_p, _dp, _dp, _dp])
_sig("oracle_validate", C.c_double, [C.c_void_p, _dp, C.c_int, _ip, _dp])
_sig("oracle_surface", C.c_int, [C.c_int, C.c_double, _dp])
_sig("oracle_regpos", C.c_int, [C.c_int, C.c_double, _dp])
_sig("oracle_pinv", C.c_int, [_dp, C.c_int, C.c_int, C.c_double, _dp])
_sig("oracle_svd_values", C.c_int, [_dp, C.c_int, C.c_int, _dp])
_sig("oracle_rfft3", C.c_int, [C.c_int, _dp, _dp])
_sig("oracle_irfft3", C.c_int, [C.c_int, _dp, _dp])
_sig("oracle_translation_spectra", C.c_int, [C.c_int, C.c_int, _dp])
_sig("oracle_direct", C.c_int, [C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp, _dp, C.c_int, _dp])
_sig("oracle_drand48", C.c_int, [C.c_long, C.c_int, _dp])
_sig("oracle_set_reverse_order", None, [C.c_int])


def dp(a):
    return a.ctypes.data_as(_dp)


def ip(a):
    return a.ctypes.data_as(_ip)


def n_e(np_):
    return np_ ** 3 - (np_ - 2) ** 3


def n_uc(np_):
    return (np_ + 2) ** 3 - np_ ** 3


class Oracle:
    def __init__(self, n, dist="uniform", q=64, np_=4, seed=26, coincide=0, points=None):
        self.n, self.np = n, np_
        if points is None:
            self.h = lib.oracle_create(n, dist.encode(), q, np_, seed, coincide)
        else:
            arrs = [np.ascontiguousarray(points[k], dtype=np.float64)
                    for k in ("sx", "sy", "sz", "den", "tx", "ty", "tz")]
            self._keep = arrs
            self.h = lib.oracle_create_from_points(n, *[dp(a) for a in arrs], q, np_)
        if not self.h:
            raise ValueError("oracle_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib.oracle_destroy(self.h)
            self.h = None

    def tree(self):
        nn = lib.oracle_num_nodes(self.h)
        a = {k: np.zeros(nn, dtype=np.int32) for k in
             ("parent", "child", "depth", "tag", "src_beg", "src_num", "trg_beg", "trg_num")}
        path = np.zeros(3 * nn, dtype=np.int32)
        lib.oracle_tree_arrays(self.h, ip(a["parent"]), ip(a["child"]), ip(a["depth"]), ip(a["tag"]),
                               ip(path), ip(a["src_beg"]), ip(a["src_num"]), ip(a["trg_beg"]), ip(a["trg_num"]))
        a["path"] = path.reshape(nn, 3)
        a["n_nodes"] = nn
        lv = np.zeros(lib.oracle_num_levels(self.h), dtype=np.int32)
        lib.oracle_level_starts(self.h, ip(lv))
        a["level_start"] = lv
        for k, L in enumerate("uvwx"):
            nnz = lib.oracle_list_nnz(self.h, k)
            ptr = np.zeros(nn + 1, dtype=np.int32)
            idx = np.zeros(max(nnz, 1), dtype=np.int32)
            lib.oracle_list_csr(self.h, k, ip(ptr), ip(idx))
            a[f"{L}_ptr"], a[f"{L}_idx"] = ptr, idx[:nnz]
        return a

    def points(self):
        out = [np.zeros(self.n) for _ in range(7)]
        lib.oracle_points(self.h, *[dp(x) for x in out])
        return dict(zip(("sx", "sy", "sz", "den", "tx", "ty", "tz"), out))

    def perms(self):
        s = np.zeros(self.n, dtype=np.int32)
        t = np.zeros(self.n, dtype=np.int32)
        lib.oracle_perms(self.h, ip(s), ip(t))
        return s, t

    def build_operators(self, mode=KERNEL_EXACT):
        lib.oracle_build_operators(self.h, mode)

    def set_operators(self, uc2ue_raw, ue2uc_raw, dc2de_raw, de2dc_raw):
        """Raw column-major arrays as laid out by kifmm_operators."""
        self._ops = [np.ascontiguousarray(x, dtype=np.float64) for x in (uc2ue_raw, ue2uc_raw, dc2de_raw, de2dc_raw)]
        lib.oracle_set_operators(self.h, *[dp(x) for x in self._ops])

    def get_operators(self):
        ne, nuc = n_e(self.np), n_uc(self.np)
        a = [np.zeros(ne * nuc), np.zeros(8 * nuc * ne), np.zeros(ne * ne), np.zeros(8 * ne * ne)]
        if lib.oracle_get_operators(self.h, *[dp(x) for x in a]) != 0:
            raise RuntimeError("operators not built")
        return a

    def evaluate(self, mode=KERNEL_EXACT, nthreads=0):
        pot = np.zeros(self.n)
        t = np.zeros(8)
        lib.oracle_evaluate(self.h, mode, nthreads, dp(pot), dp(t))
        return pot, dict(zip(("up", "ulist", "vlist", "wlist", "xlist", "down", "copy", "total"), t.tolist()))

    def intermediates(self):
        nn = lib.oracle_num_nodes(self.h)
        ne = n_e(self.np)
        a = [np.zeros(nn * ne) for _ in range(3)]
        lib.oracle_intermediates(self.h, *[dp(x) for x in a])
        return [x.reshape(nn, ne) for x in a]

    def validate(self, pot_orig, nv=1000):
        pot_orig = np.ascontiguousarray(pot_orig, dtype=np.float64)
        ids = np.zeros(nv, dtype=np.int32)
        direct = np.zeros(nv)
        r = lib.oracle_validate(self.h, dp(pot_orig), nv, ip(ids), dp(direct))
        return r, ids, direct


def surface(np_, R):
    n = n_e(np_)
    xyz = np.zeros(3 * n)
    lib.oracle_surface(np_, R, dp(xyz))
    return xyz.reshape(3, n)


def regpos(np_, R=1.0):
    n = (2 * np_) ** 3
    xyz = np.zeros(3 * n)
    lib.oracle_regpos(np_, R, dp(xyz))
    return xyz.reshape(3, n)


def pinv(M, eps):
    M = np.asfortranarray(M, dtype=np.float64)
    m, n = M.shape
    R = np.zeros((n, m), order="F")
    lib.oracle_pinv(dp(M), m, n, eps, dp(R))
    return R


def svd_values(M):
    M = np.asfortranarray(M, dtype=np.float64)
    m, n = M.shape
    s = np.zeros(min(m, n))
    lib.oracle_svd_values(dp(M), m, n, dp(s))
    return s


def rfft3(g):
    N = g.shape[0]
    g = np.ascontiguousarray(g, dtype=np.float64)
    out = np.zeros((N, N, N // 2 + 1), dtype=np.complex128)
    lib.oracle_rfft3(N, dp(g), out.ctypes.data_as(_dp))
    return out


def irfft3(Y):
    N = Y.shape[0]
    Y = np.ascontiguousarray(Y, dtype=np.complex128)
    out = np.zeros((N, N, N))
    lib.oracle_irfft3(N, Y.ctypes.data_as(_dp), dp(out))
    return out


def translation_spectra(np_, mode=KERNEL_EXACT):
    N = 2 * np_
    out = np.zeros((343, N, N, N // 2 + 1), dtype=np.complex128)
    lib.oracle_translation_spectra(np_, mode, out.ctypes.data_as(_dp))
    return out


def direct(t, s, w, mode=KERNEL_EXACT):
    t = [np.ascontiguousarray(x, dtype=np.float64) for x in t]
    s = [np.ascontiguousarray(x, dtype=np.float64) for x in s]
    w = np.ascontiguousarray(w, dtype=np.float64)
    pot = np.zeros(len(t[0]))
    lib.oracle_direct(len(t[0]), *[dp(x) for x in t], len(s[0]), *[dp(x) for x in s], dp(w), mode, dp(pot))
    return pot


def drand48(seed, k):
    out = np.zeros(k)
    lib.oracle_drand48(seed, k, dp(out))
    return out


def set_reverse_order(on):
    lib.oracle_set_reverse_order(1 if on else 0)


# Parity bound per NP.  NP <= 4: 1e-12.  NP = 6: the UC2UE / DC2DE
# pseudo-inverses keep singular values down to 1e-7 of the largest
# (cutoff 0.1^(NP+1), fmm/trans.cc:120-124), so rounding-level changes in
# summation order move the potentials by ~2e-11 relative; the oracle
# against itself with only its source sums reversed differs by 1.6e-11 to
# 2.0e-11 (test_oracle.py::test_summation_order_floor).  The bound is 3x
# that floor.
PARITY_RTOL = {3: 1e-12, 4: 1e-12, 6: 6e-11}
