"""Learned cost model and tuner: picks points per box and the precision path.

A 16 -> 64 -> 64 -> 4 MLP (FP32, include/kifmm_mlp.h) maps 16 statistics
of the octree an input produces at a given points-per-box q to the log
phase times (up, U, V, down) of one evaluation.  `Tuner.choose` builds the
tree for every candidate q (host work only), predicts the total time of
every allowed precision path and returns the fastest; without a model it
falls back to an exhaustive sweep that times every candidate on the GPU.
The network's forward, backward and Adam steps run as HIP kernels
(csrc/hip/kifmm_mlp.hip); this module holds the host logic around them.
No counterpart in the reference: the nearest knobs are its hand-set BS /
UPBS blocking (fmm/util.c:118-164) and the q on its command line
(fmm/driver.cc:211-240).
"""
from __future__ import annotations

import ctypes as C
import json
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

import kifmm_amd as K

NIN, NHID, NOUT = 16, 64, 4
NPARAMS = NIN * NHID + NHID + NHID * NHID + NHID + NHID * NOUT + NOUT
Q_CHOICES = (32, 64, 128, 256, 512, 1024)
PHASES = ("up", "ulist", "vlist", "down")
FEATURES = ("log2 sources", "log2 nodes", "max depth", "log2 leaves", "log2 U pairs", "log2 V entries",
            "log2 (1 + W pairs)", "log2 (1 + X pairs)", "log2 S2M pairs", "log2 L2T pairs", "log2 GEMM flops",
            "NP", "mixed precision", "log2 q", "leaf occupancy CV", "deepest-level leaf fraction")
WEIGHTS = Path(__file__).resolve().parent / "costmodel_weights.json"

_fp = C.POINTER(C.c_float)
for _name, _args in (("kifmm_mlp_create", [C.c_int, _fp, C.POINTER(C.c_void_p)]),
                     ("kifmm_mlp_destroy", [C.c_void_p]),
                     ("kifmm_mlp_get_params", [C.c_void_p, _fp]),
                     ("kifmm_mlp_predict", [C.c_void_p, _fp, C.c_int, _fp]),
                     ("kifmm_mlp_gradient", [C.c_void_p, _fp, _fp, C.c_int, _fp, C.POINTER(C.c_double)]),
                     ("kifmm_mlp_train", [C.c_void_p, _fp, _fp, C.c_int, C.c_int, C.c_float,
                                          C.POINTER(C.c_double)])):
    _f = getattr(K.lib, _name)
    _f.restype = C.c_int
    _f.argtypes = _args


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


def features(problem: K.Problem, q: int, precision: int, operators=None) -> np.ndarray:
    """The 16 statistics (FEATURES) of `problem`'s tree at points per box q
    for a precision path; work counts come from the evaluator's own host
    planner (kifmm_gpu_plan), so they are the counts the GPU executes.
    `operators` (same NP) avoids building them for this problem."""
    t = problem.tree()
    w = K.gpu_plan(problem, operators=operators)
    leaf = t["child"] == -1
    occ = t["src_num"][leaf & (t["src_num"] > 0)].astype(np.float64)
    depth = t["depth"]
    dmax = int(depth.max())
    lg = lambda x: float(np.log2(max(float(x), 1.0)))
    f = [lg(t["sx"].size), lg(t["n_nodes"]), float(dmax), lg(leaf.sum()), lg(w["u_pairs"]), lg(w["v_entries"]),
         lg(1 + w["w_pairs"]), lg(1 + w["x_pairs"]), lg(w["s2m_pairs"]), lg(w["l2t_pairs"]), lg(w["gemm_flops"]),
         float(problem.view().np), float(precision == K.PREC_MIXED), lg(q),
         float(occ.std() / occ.mean()) if occ.size else 0.0,
         float((leaf & (depth == dmax)).sum() / max(leaf.sum(), 1))]
    return np.asarray(f, dtype=np.float32)


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for din, dout in ((NIN, NHID), (NHID, NHID), (NHID, NOUT)):
        lim = np.sqrt(6.0 / din)
        parts += [rng.uniform(-lim, lim, din * dout), np.zeros(dout)]
    return np.concatenate(parts).astype(np.float32)


class Network:
    """Device-resident network (kifmm_mlp handle)."""

    def __init__(self, params: np.ndarray, device: int = 0):
        p, pp = _f32(params)
        if p.size != NPARAMS:
            raise ValueError(f"expected {NPARAMS} parameters, got {p.size}")
        h = C.c_void_p()
        K._check(K.lib.kifmm_mlp_create(device, pp, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            K.lib.kifmm_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(NPARAMS, dtype=np.float32)
        K._check(K.lib.kifmm_mlp_get_params(self._h, out.ctypes.data_as(_fp)))
        return out

    def predict(self, X: np.ndarray) -> np.ndarray:
        X, xp = _f32(X)
        Y = np.zeros((X.shape[0], NOUT), dtype=np.float32)
        K._check(K.lib.kifmm_mlp_predict(self._h, xp, X.shape[0], Y.ctypes.data_as(_fp)))
        return Y

    def gradient(self, X: np.ndarray, T: np.ndarray):
        X, xp = _f32(X)
        T, tp = _f32(T)
        g = np.zeros(NPARAMS, dtype=np.float32)
        loss = C.c_double()
        K._check(K.lib.kifmm_mlp_gradient(self._h, xp, tp, X.shape[0], g.ctypes.data_as(_fp), C.byref(loss)))
        return g, loss.value

    def train(self, X: np.ndarray, T: np.ndarray, epochs: int, lr: float) -> np.ndarray:
        X, xp = _f32(X)
        T, tp = _f32(T)
        losses = np.zeros(epochs)
        K._check(K.lib.kifmm_mlp_train(self._h, xp, tp, X.shape[0], epochs, lr,
                                       losses.ctypes.data_as(C.POINTER(C.c_double))))
        return losses


@dataclass
class CostModel:
    """Trained network plus the feature / target standardisation."""
    params: np.ndarray
    x_mean: np.ndarray
    x_std: np.ndarray
    y_mean: np.ndarray
    y_std: np.ndarray
    meta: dict = field(default_factory=dict)

    @staticmethod
    def fit(X: np.ndarray, times_ms: np.ndarray, epochs: int = 3000, lr: float = 3e-3, seed: int = 0,
            device: int = 0) -> "CostModel":
        """Train on rows of features and measured phase times (ms), the
        targets being standardised log times."""
        X = np.asarray(X, dtype=np.float64)
        Y = np.log(np.maximum(np.asarray(times_ms, dtype=np.float64), 1e-3))
        xm, xs = X.mean(0), X.std(0) + 1e-6
        ym, ys = Y.mean(0), Y.std(0) + 1e-6
        net = Network(init_params(seed), device)
        losses = net.train((X - xm) / xs, (Y - ym) / ys, epochs, lr)
        m = CostModel(net.params(), xm, xs, ym, ys, {"epochs": epochs, "lr": lr, "final_loss": float(losses[-1]),
                                                   "first_loss": float(losses[0]), "rows": int(X.shape[0])})
        net.close()
        return m

    def predict_ms(self, X: np.ndarray, device: int = 0) -> np.ndarray:
        """Phase times (ms) predicted for rows of features."""
        net = Network(self.params, device)
        z = net.predict((np.atleast_2d(X) - self.x_mean) / self.x_std).astype(np.float64)
        net.close()
        return np.exp(z * self.y_std + self.y_mean)

    def save(self, path: Path = WEIGHTS):
        d = {"params": self.params.tolist(), "x_mean": self.x_mean.tolist(), "x_std": self.x_std.tolist(),
             "y_mean": self.y_mean.tolist(), "y_std": self.y_std.tolist(), "features": list(FEATURES),
             "targets": ["log ms " + p for p in PHASES], "meta": self.meta}
        Path(path).write_text(json.dumps(d))

    @staticmethod
    def load(path: Path = WEIGHTS) -> "CostModel | None":
        p = Path(path)
        if not p.exists():
            return None
        d = json.loads(p.read_text())
        arr = lambda k, t=np.float64: np.asarray(d[k], dtype=t)
        return CostModel(arr("params", np.float32), arr("x_mean"), arr("x_std"), arr("y_mean"), arr("y_std"),
                         d.get("meta", {}))


def phase_times(problem: K.Problem, precision: int, steps: int = 3, warmup: int = 1, graph: bool = False,
                slow_ms: float = 50.0):
    """Measured mean phase times (ms) and wall ms per evaluation.  An
    evaluation slower than slow_ms is measured once (its first run)."""
    ev = K.GpuEvaluator(problem, precision=precision, graph=graph)
    acc = np.zeros(len(PHASES))
    t0 = time.perf_counter()
    _, t = ev.evaluate(copy=False)
    first = (time.perf_counter() - t0) * 1e3
    if first > slow_ms:
        ev.close()
        return np.array([t[p] for p in PHASES]), first
    for _ in range(max(warmup - 1, 0)):
        ev.evaluate(copy=False)
    t0 = time.perf_counter()
    for _ in range(steps):
        _, t = ev.evaluate(copy=False)
        acc += [t[p] for p in PHASES]
    wall = (time.perf_counter() - t0) / steps * 1e3
    ev.close()
    return acc / steps, wall


def paths_for(np_: int):
    return (K.PREC_FP64, K.PREC_MIXED) if np_ <= 4 else (K.PREC_FP64,)


class Tuner:
    """Choose (q, precision) for one input: by the cost model when a model
    is available, else by timing every candidate (exhaustive sweep).  With
    verify = k > 1 the model ranks the candidates and the k best-ranked are
    timed on the GPU, the fastest measured one kept: a guard against a
    misranked near-tie at the cost of k tree builds and a few evaluations."""

    def __init__(self, model: CostModel | None = None, qs=Q_CHOICES, allow_mixed: bool = True):
        self.model = model
        self.qs = tuple(qs)
        self.allow_mixed = allow_mixed
        self.model_pick = None  # (q, precision, predicted ms) of the last model ranking

    def _problem(self, pts, q, np_):
        return K.Problem.from_points(*pts, pts_max=q, np_=np_)

    def choose(self, pts, np_: int, verify: int = 1):
        """pts = (sx, sy, sz, den, tx, ty, tz) in the root box.  Returns
        (q, precision, predicted or measured ms, method)."""
        paths = paths_for(np_) if self.allow_mixed else (K.PREC_FP64,)
        best = None
        if self.model is not None:
            rows, keys, ops = [], [], None
            for q in self.qs:
                prob = self._problem(pts, q, np_)
                if ops is None:
                    ops = prob.operators()
                for pr in paths:
                    rows.append(features(prob, q, pr, ops))
                    keys.append((q, pr))
            tot = self.model.predict_ms(np.stack(rows)).sum(1)
            i0 = int(np.argmin(tot))
            self.model_pick = (keys[i0][0], keys[i0][1], float(tot[i0]))  # the network's own choice
            if verify <= 1:
                i = int(np.argmin(tot))
                return keys[i][0], keys[i][1], float(tot[i]), "model"
            for i in np.argsort(tot, kind="stable")[:verify]:
                q, pr = keys[int(i)]
                prob = self._problem(pts, q, np_)
                prob.operators()
                _, wall = phase_times(prob, pr)
                if best is None or wall < best[2]:
                    best = (q, pr, wall)
            return best[0], best[1], best[2], "model+verify"
        for q in self.qs:
            prob = self._problem(pts, q, np_)
            prob.operators()
            for pr in paths:
                _, wall = phase_times(prob, pr)
                if best is None or wall < best[2]:
                    best = (q, pr, wall)
        return best[0], best[1], best[2], "sweep"
