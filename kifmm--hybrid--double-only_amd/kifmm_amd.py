"""ctypes bindings of libkifmm (the C ABI in include/kifmm_host.h and include/kifmm_hip.h).

This mirrors the reference's evaluator plug-in surface (fmm/evaluate.h:16-57):
``Problem`` stands for the host setup that main() performs before ``preproc``
(fmm/driver.cc:244-329), ``GpuEvaluator.create`` for ``preproc``
(fmm/evaluate--cuda.cu:1441) and ``GpuEvaluator.evaluate`` for ``run``
(fmm/run--cuda.cu:27).  Errors raise ``KifmmError`` carrying the library's
``kifmm_last_error()`` text; nothing falls back to a CPU path: if the shared
library is missing the import fails.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "libkifmm.so"

KIFMM_SRCNODE = 1
KIFMM_TRGNODE = 2
ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EDEVICE", -4: "EUNSUP"}
PREC_FP64, PREC_MIXED = 0, 1
GPU_GRAPH = 1
GPU_NO_CLOSE_LIST = 2
GPU_MIXED_INLINE = GPU_NO_CLOSE_LIST  # earlier name of the same flag
GPU_OVERLAP = 4
PHASE_UP, PHASE_U, PHASE_V, PHASE_W, PHASE_X, PHASE_DOWN, PHASE_ALL = 1, 2, 4, 8, 16, 32, 63
LEVEL_AUTO, LEVEL_GEMM, LEVEL_FUSED = 0, 1, 2


class KifmmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


if not LIB_PATH.exists():
    raise ImportError(f"libkifmm not built: {LIB_PATH} is missing (run __graft_entry__.build())")
lib = C.CDLL(str(LIB_PATH))

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


class TreeView(C.Structure):
    _fields_ = [
        ("np", C.c_int), ("n_nodes", C.c_int), ("n_levels", C.c_int), ("level_start", _ip),
        ("parent", _ip), ("child", _ip), ("depth", _ip), ("tag", _ip), ("path", _ip),
        ("src_beg", _ip), ("src_num", _ip), ("trg_beg", _ip), ("trg_num", _ip),
        ("u_ptr", _ip), ("u_idx", _ip), ("v_ptr", _ip), ("v_idx", _ip),
        ("w_ptr", _ip), ("w_idx", _ip), ("x_ptr", _ip), ("x_idx", _ip),
        ("n_src", C.c_int), ("n_trg", C.c_int),
        ("sx", _dp), ("sy", _dp), ("sz", _dp), ("sden", _dp),
        ("tx", _dp), ("ty", _dp), ("tz", _dp),
        ("coincident", C.c_int),
    ]


class Operators(C.Structure):
    _fields_ = [("np", C.c_int), ("n_e", C.c_int), ("n_uc", C.c_int),
                ("uc2ue", _dp), ("ue2uc", _dp), ("dc2de", _dp), ("de2dc", _dp)]


class ProblemOpts(C.Structure):
    _fields_ = [("np", C.c_int), ("seed", C.c_long), ("coincide", C.c_int), ("nthreads", C.c_int)]


class GpuOptions(C.Structure):
    _fields_ = [("device", C.c_int), ("precision", C.c_int), ("flags", C.c_int), ("reserved", C.c_int * 5)]


class GpuTuning(C.Structure):
    _fields_ = [("p2p_form", C.c_int), ("hadamard_fb", C.c_int), ("hadamard_form", C.c_int),
                ("overlap", C.c_int), ("up_form", C.c_int), ("down_form", C.c_int), ("reserved", C.c_int * 2)]


P2P_STATIC, P2P_PERSISTENT, P2P_STATIC64, P2P_SPLIT = 0, 1, 2, 3
HAD_VALU, HAD_MFMA = 0, 1


class PhaseTimes(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "up", "ulist", "vlist", "v_fft_src", "v_trans", "v_hadamard", "v_ifft", "wlist", "xlist",
        "down", "copy", "total", "p2p_kernel")] + [("reserved", C.c_double * 3)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "reserved"}


class WorkCounts(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "u_pairs", "s2m_pairs", "l2t_pairs", "w_pairs", "x_pairs", "v_entries", "v_cmac", "v_cmac_ref",
        "gemm_flops", "fft_bytes", "hadamard_bytes", "close_pairs", "copy_chunks", "pair_items")] + \
        [("reserved", C.c_double * 2)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "reserved"}


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("kifmm_last_error", C.c_char_p, [])
_sig("kifmm_problem_create", C.c_void_p, [C.c_int, C.c_char_p, C.c_int, C.POINTER(ProblemOpts)])
_sig("kifmm_problem_from_points", C.c_void_p, [C.c_int] + [_dp] * 7 + [C.c_int, C.c_int])
_sig("kifmm_problem_create_gpu", C.c_void_p, [C.c_int, C.c_int, C.c_char_p, C.c_int, C.POINTER(ProblemOpts)])
_sig("kifmm_problem_from_points_gpu", C.c_void_p, [C.c_int, C.c_int] + [_dp] * 7 + [C.c_int, C.c_int])
_sig("kifmm_problem_destroy", None, [C.c_void_p])
_sig("kifmm_problem_view", C.c_int, [C.c_void_p, C.POINTER(TreeView)])
_sig("kifmm_problem_operators", C.c_int, [C.c_void_p, C.POINTER(Operators)])
_sig("kifmm_problem_scatter", C.c_int, [C.c_void_p, _dp, _dp])
_sig("kifmm_problem_points", C.c_int, [C.c_void_p] + [_dp] * 7)
_sig("kifmm_problem_validate", C.c_double, [C.c_void_p, _dp, C.c_int, _ip, _dp])
_sig("kifmm_points_write", C.c_int, [C.c_char_p, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp])
_sig("kifmm_problem_from_file", C.c_void_p, [C.c_char_p, C.c_int, C.c_int])
_sig("kifmm_problem_validate_gpu", C.c_double, [C.c_void_p, C.c_void_p, _dp, C.c_int, _ip, _dp])
_sig("kifmm_gpu_direct", C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp])
_sig("kifmm_problem_setup_times", C.c_int, [C.c_void_p, _dp])
for _n in ("accuracy", "validate", "coincide", "block_size", "upbs"):
    _sig(f"kifmm_getenv_{_n}", C.c_int, [])
_sig("kifmm_getenv_seed", C.c_long, [])
_sig("kifmm_surface", C.c_int, [C.c_int, C.c_double, _dp])
_sig("kifmm_pinv", C.c_int, [_dp, C.c_int, C.c_int, C.c_double, _dp])
_sig("kifmm_drand48", C.c_int, [C.c_long, C.c_int, _dp])
_sig("kifmm_gpu_available", C.c_int, [])
_sig("kifmm_gpu_create", C.c_int, [C.POINTER(TreeView), C.POINTER(Operators), C.POINTER(GpuOptions),
                                   C.POINTER(C.c_void_p)])
_sig("kifmm_gpu_evaluate", C.c_int, [C.c_void_p, _dp, C.POINTER(PhaseTimes)])
_sig("kifmm_gpu_destroy", C.c_int, [C.c_void_p])
_sig("kifmm_gpu_evaluate_phases", C.c_int, [C.c_void_p, C.c_uint, _dp, C.POINTER(PhaseTimes)])
_sig("kifmm_gpu_work_counts", C.c_int, [C.c_void_p, C.POINTER(WorkCounts)])
_sig("kifmm_gpu_copy_intermediate", C.c_int, [C.c_void_p, C.c_int, _dp, C.c_long])
_sig("kifmm_gpu_plan", C.c_int, [C.POINTER(TreeView), C.POINTER(Operators), C.POINTER(WorkCounts)])
_sig("kifmm_gpu_set_tuning", C.c_int, [C.c_void_p, C.POINTER(GpuTuning)])
_sig("kifmm_gpu_get_tuning", C.c_int, [C.c_void_p, C.POINTER(GpuTuning)])
_sig("kifmm_gpu_autotune", C.c_int, [C.c_void_p, C.c_int, C.POINTER(GpuTuning), _dp, C.c_int])
_sig("kifmm_gpu_output", _dp, [C.c_void_p])
_sig("kifmm_gpu_direct_noalloc", C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp])
_sig("kifmm_gpu_set_densities", C.c_int, [C.c_void_p, _dp])
_sig("kifmm_gpu_create_times", C.c_int, [C.c_void_p, _dp, C.c_int])
_sig("kifmm_gpu_close_list", C.c_long, [C.c_void_p, _ip, C.c_long, _ip, C.c_long])


def last_error() -> str:
    return (lib.kifmm_last_error() or b"").decode()


def _check(rc):
    if rc < 0:
        raise KifmmError(rc, last_error())
    return rc


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def n_e(np_):
    return np_ ** 3 - (np_ - 2) ** 3


def n_uc(np_):
    return (np_ + 2) ** 3 - np_ ** 3


class Problem:
    """Points, octree, U/V/W/X lists and gathered arrays (host)."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def create(cls, n, distribution="uniform", pts_max=64, np_=None, seed=None, coincide=None, nthreads=0,
               device=None):
        """device=None builds the tree and lists on the host, an int on that GPU (same arrays)."""
        opts = ProblemOpts(np_ if np_ else -1, -1 if seed is None else seed,
                           -1 if coincide is None else int(coincide), nthreads)
        if device is None:
            h = lib.kifmm_problem_create(n, distribution.encode(), pts_max, C.byref(opts))
        else:
            h = lib.kifmm_problem_create_gpu(int(device), n, distribution.encode(), pts_max, C.byref(opts))
        if not h:
            raise KifmmError(-1, last_error())
        return cls(h)

    @classmethod
    def from_points(cls, sx, sy, sz, den, tx=None, ty=None, tz=None, pts_max=64, np_=6, device=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (sx, sy, sz, den)]
        if tx is None:
            t = arrs[:3]
        else:
            t = [np.ascontiguousarray(a, dtype=np.float64) for a in (tx, ty, tz)]
        ptrs = [_dptr(a) for a in arrs + t]
        if device is None:
            h = lib.kifmm_problem_from_points(len(arrs[0]), *ptrs, pts_max, np_)
        else:
            h = lib.kifmm_problem_from_points_gpu(int(device), len(arrs[0]), *ptrs, pts_max, np_)
        if not h:
            raise KifmmError(-1, last_error())
        p = cls(h)
        p._keep = arrs + t
        return p

    def close(self):
        if self._h:
            lib.kifmm_problem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def view(self) -> TreeView:
        v = TreeView()
        _check(lib.kifmm_problem_view(self._h, C.byref(v)))
        return v

    def operators(self) -> Operators:
        o = Operators()
        _check(lib.kifmm_problem_operators(self._h, C.byref(o)))
        return o

    def operator_arrays(self):
        o = self.operators()
        ne, nuc = o.n_e, o.n_uc
        return {
            "uc2ue": _arr(o.uc2ue, ne * nuc, np.float64).reshape(nuc, ne).T,
            "ue2uc": _arr(o.ue2uc, 8 * nuc * ne, np.float64).reshape(8, ne, nuc).transpose(0, 2, 1),
            "dc2de": _arr(o.dc2de, ne * ne, np.float64).reshape(ne, ne).T,
            "de2dc": _arr(o.de2dc, 8 * ne * ne, np.float64).reshape(8, ne, ne).transpose(0, 2, 1),
            "raw": o,
        }

    def tree(self):
        """Numpy copies of the tree view (node arrays, CSR lists, gathered points)."""
        v = self.view()
        nn = v.n_nodes
        t = {"np": v.np, "n_nodes": nn, "coincident": v.coincident,
             "level_start": _arr(v.level_start, v.n_levels, np.int32)}
        for k in ("parent", "child", "depth", "tag", "src_beg", "src_num", "trg_beg", "trg_num"):
            t[k] = _arr(getattr(v, k), nn, np.int32)
        t["path"] = _arr(v.path, 3 * nn, np.int32).reshape(nn, 3)
        for L in "uvwx":
            ptr = _arr(getattr(v, f"{L}_ptr"), nn + 1, np.int32)
            t[f"{L}_ptr"] = ptr
            t[f"{L}_idx"] = _arr(getattr(v, f"{L}_idx"), int(ptr[-1]), np.int32)
        for k in ("sx", "sy", "sz", "sden"):
            t[k] = _arr(getattr(v, k), v.n_src, np.float64)
        for k in ("tx", "ty", "tz"):
            t[k] = _arr(getattr(v, k), v.n_trg, np.float64)
        return t

    def points(self):
        v = self.view()
        n = v.n_src
        out = [np.zeros(n) for _ in range(7)]
        _check(lib.kifmm_problem_points(self._h, *[_dptr(a) for a in out]))
        return dict(zip(("sx", "sy", "sz", "den", "tx", "ty", "tz"), out))

    def scatter(self, pot_leaf):
        pot_leaf = np.ascontiguousarray(pot_leaf, dtype=np.float64)
        out = np.zeros_like(pot_leaf)
        _check(lib.kifmm_problem_scatter(self._h, _dptr(pot_leaf), _dptr(out)))
        return out

    def validate(self, pot_orig, nv=1000):
        pot_orig = np.ascontiguousarray(pot_orig, dtype=np.float64)
        ids = np.zeros(nv, dtype=np.int32)
        direct = np.zeros(nv)
        r = lib.kifmm_problem_validate(self._h, _dptr(pot_orig), nv, ids.ctypes.data_as(_ip), _dptr(direct))
        if r < 0:
            raise KifmmError(-1, last_error())
        return r, ids, direct

    @classmethod
    def from_file(cls, path, pts_max=64, np_=6):
        """Points from a KIFMMPT1 file (kifmm_problem_from_file)."""
        h = lib.kifmm_problem_from_file(str(path).encode(), pts_max, np_)
        if not h:
            raise KifmmError(-1, last_error())
        return cls(h)

    def validate_gpu(self, evaluator, pot_orig, nv=1000):
        """validate() with the direct sums on the GPU (kifmm_problem_validate_gpu):
        same sample, same error definition; `evaluator` holds this problem."""
        pot_orig = np.ascontiguousarray(pot_orig, dtype=np.float64)
        ids = np.zeros(nv, dtype=np.int32)
        direct = np.zeros(nv)
        r = lib.kifmm_problem_validate_gpu(self._h, evaluator._h, _dptr(pot_orig), nv, ids.ctypes.data_as(_ip),
                                           _dptr(direct))
        if r < 0:
            raise KifmmError(-1, last_error())
        return r, ids, direct

    def setup_times(self):
        t = np.zeros(3)
        _check(lib.kifmm_problem_setup_times(self._h, _dptr(t)))
        return dict(zip(("generate", "tree_lists", "operators"), t.tolist()))


def gpu_plan(problem: Problem, view: TreeView | None = None, operators: Operators | None = None):
    """Host-only derivation and bounds check of the device work lists
    (operators only fix NP and the surface sizes here: any set of the same
    NP serves)."""
    v = view if view is not None else problem.view()
    w = WorkCounts()
    ops = operators if operators is not None else problem.operators()
    _check(lib.kifmm_gpu_plan(C.byref(v), C.byref(ops), C.byref(w)))
    return w.as_dict()


def write_points(path, sx, sy, sz, den, tx=None, ty=None, tz=None):
    """Write a KIFMMPT1 point file (kifmm_points_write); no targets = the
    sources are the targets."""
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (sx, sy, sz, den)]
    t = [None] * 3 if tx is None else [np.ascontiguousarray(v, dtype=np.float64) for v in (tx, ty, tz)]
    _check(lib.kifmm_points_write(str(path).encode(), len(a[0]), *[_dptr(v) for v in a],
                                  *[(_dptr(v) if v is not None else None) for v in t]))


def gpu_available() -> bool:
    return bool(lib.kifmm_gpu_available())


class GpuEvaluator:
    """Device copy of one problem (kifmm_gpu_create) and its evaluation."""

    def __init__(self, problem: Problem, device=0, precision=PREC_FP64, operators: Operators | None = None,
                 graph=False, inline_close=False, overlap=False):
        """inline_close: do not list the close U-list pairs at create time;
        the mixed path then tests for them inside its loop and the FP64 path
        forms r^2 from coordinate differences instead of the Gram form.
        overlap: run the U-list on a second stream alongside up and V."""
        self.problem = problem
        self._view = problem.view()
        self._ops = operators if operators is not None else problem.operators()
        opts = GpuOptions(device, precision, (GPU_GRAPH if graph else 0) | (GPU_NO_CLOSE_LIST if inline_close else 0)
                          | (GPU_OVERLAP if overlap else 0))
        h = C.c_void_p()
        _check(lib.kifmm_gpu_create(C.byref(self._view), C.byref(self._ops), C.byref(opts), C.byref(h)))
        self._h = h
        self.n_trg = self._view.n_trg
        self.n_nodes = self._view.n_nodes
        self.np = self._view.np

    def close(self):
        if getattr(self, "_h", None):
            self._out = None  # the pinned buffer goes with the handle
            lib.kifmm_gpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def output(self):
        """The handle's pinned host buffer of potentials (gathered order) as a
        numpy view, overwritten by every evaluate(copy="view"); None without
        pinned memory."""
        if getattr(self, "_out", None) is None:
            p = lib.kifmm_gpu_output(self._h)
            self._out = np.ctypeslib.as_array(p, shape=(self.n_trg,)) if p and self.n_trg else None
        return self._out

    def evaluate(self, copy=True):
        """Returns (potentials in gathered order or None, phase times in ms).
        copy=True: the potentials come back into the handle's pinned buffer
        (chunk by chunk while the U-list still runs) and a numpy copy of them
        is returned; copy="view": the pinned buffer itself is returned (no
        host pass; the next evaluation overwrites it); copy=False: they stay
        on the device."""
        t = PhaseTimes()
        out = self.output() if copy else None
        if copy and out is None:  # no pinned memory: straight into a numpy array
            pot = np.zeros(self.n_trg)
            _check(lib.kifmm_gpu_evaluate(self._h, _dptr(pot), C.byref(t)))
            return pot, t.as_dict()
        _check(lib.kifmm_gpu_evaluate(self._h, _dptr(out) if copy else None, C.byref(t)))
        if not copy:
            return None, t.as_dict()
        return (out if copy == "view" else out.copy()), t.as_dict()

    def evaluate_into(self, pot):
        """kifmm_gpu_evaluate into a caller array (float64, n_trg): the pinned
        copy, then one host pass into it."""
        assert pot.dtype == np.float64 and pot.flags.c_contiguous and pot.size == self.n_trg
        t = PhaseTimes()
        _check(lib.kifmm_gpu_evaluate(self._h, _dptr(pot), C.byref(t)))
        return t.as_dict()

    def direct(self, tx, ty, tz, noalloc=False):
        """FP64 direct sums over this handle's sources at the given points
        (kifmm_gpu_direct, or kifmm_gpu_direct_noalloc with noalloc=True)."""
        t = [np.ascontiguousarray(a, dtype=np.float64) for a in (tx, ty, tz)]
        out = np.zeros(t[0].size)
        f = lib.kifmm_gpu_direct_noalloc if noalloc else lib.kifmm_gpu_direct
        _check(f(self._h, t[0].size, *[_dptr(a) for a in t], _dptr(out)))
        return out

    def set_densities(self, den_leaf):
        """New source densities in gathered source order (kifmm_gpu_set_densities)."""
        den = np.ascontiguousarray(den_leaf, dtype=np.float64)
        assert den.size == self._view.n_src
        _check(lib.kifmm_gpu_set_densities(self._h, _dptr(den)))

    def create_times(self):
        """Wall ms of kifmm_gpu_create: host plan, upload, close-pair search,
        other device set-up, whole call."""
        t = np.zeros(5)
        _check(lib.kifmm_gpu_create_times(self._h, _dptr(t), 5))
        return dict(zip(("plan", "upload", "close_pairs", "device_setup", "total"), t.tolist()))

    def close_list(self):
        """(offsets [n_trg + 1], source indices) of the close-pair list, or None."""
        n = lib.kifmm_gpu_close_list(self._h, None, 0, None, 0)
        if n < 0:
            if n == -4:
                return None
            _check(n)
        ptr = np.zeros(self.n_trg + 1, dtype=np.int32)
        idx = np.zeros(max(n, 1), dtype=np.int32)
        _check(lib.kifmm_gpu_close_list(self._h, ptr.ctypes.data_as(_ip), ptr.size, idx.ctypes.data_as(_ip), n))
        return ptr, idx[:n]

    def evaluate_phases(self, phases, copy=True):
        t = PhaseTimes()
        out = self.output() if copy else None
        pot = (np.zeros(self.n_trg) if out is None else out) if copy else None
        _check(lib.kifmm_gpu_evaluate_phases(self._h, phases, _dptr(pot) if copy else None, C.byref(t)))
        return (pot.copy() if copy and out is not None else pot), t.as_dict()

    def tuning(self):
        t = GpuTuning()
        _check(lib.kifmm_gpu_get_tuning(self._h, C.byref(t)))
        return {"p2p_form": t.p2p_form, "hadamard_fb": t.hadamard_fb, "hadamard_form": t.hadamard_form,
                "overlap": t.overlap, "up_form": t.up_form, "down_form": t.down_form}

    def set_tuning(self, p2p_form=P2P_STATIC, hadamard_fb=8, hadamard_form=HAD_MFMA, overlap=None, up_form=None,
                   down_form=None):
        """overlap / up_form / down_form None keep the handle's current
        stream layout / level forms."""
        cur = self.tuning()
        overlap = cur["overlap"] if overlap is None else overlap
        up_form = cur["up_form"] if up_form is None else up_form
        down_form = cur["down_form"] if down_form is None else down_form
        t = GpuTuning(p2p_form, hadamard_fb, hadamard_form, int(overlap), int(up_form), int(down_form))
        _check(lib.kifmm_gpu_set_tuning(self._h, C.byref(t)))

    def autotune(self, reps=3):
        """Time the launch shapes on this input, keep the fastest; returns
        (choice, ms) with ms for STATIC, PERSISTENT, STATIC64, then the
        Hadamard on the VALU with 4, 8, 16 frequencies per block and on MFMA
        with 4, 8, 16, then whole evaluations with the copy back in stream
        layouts 0 and 1, then SPLIT (mixed path), then layouts 2 and 3, then
        up as column GEMMs and fused, down as column GEMMs and fused, then
        layouts 4, 5, 6; -1 where not timed."""
        t = GpuTuning()
        ms = np.zeros(21)
        _check(lib.kifmm_gpu_autotune(self._h, reps, C.byref(t), _dptr(ms), 21))
        return {"p2p_form": t.p2p_form, "hadamard_fb": t.hadamard_fb, "hadamard_form": t.hadamard_form,
                "overlap": t.overlap, "up_form": t.up_form, "down_form": t.down_form}, ms

    def work_counts(self):
        w = WorkCounts()
        _check(lib.kifmm_gpu_work_counts(self._h, C.byref(w)))
        return w.as_dict()

    def intermediate(self, which):
        """which: 'uden', 'dchk', 'dden' -> [n_nodes, n_e]; 'pot' -> [n_trg]."""
        idx = {"uden": 0, "dchk": 1, "dden": 2, "pot": 3}[which]
        ne = n_e(self.np)
        lde = (ne + 7) // 8 * 8
        count = self.n_trg if idx == 3 else self.n_nodes * lde
        out = np.zeros(count)
        _check(lib.kifmm_gpu_copy_intermediate(self._h, idx, _dptr(out), count))
        if idx == 3:
            return out
        return out.reshape(self.n_nodes, lde)[:, :ne]
