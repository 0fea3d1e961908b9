"""GPU path at the benchmark sizes, through properties that do not need the
CPU oracle (too slow there): linearity in the densities, the sampled error
against the FP64 direct sum at NV = 1000 (direct sums on the GPU, same
sample as validate, fmm/driver.cc:141-206), graph replay = eager, and the
mixed path's error against the FP64 path's on the same tree."""
import numpy as np
import pytest

import kifmm_amd as K
import oracle_lib as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_uniform_4m_np6_linearity_and_error():
    """4,194,304 uniform points, seed 1, NP = 6, 512 per box."""
    print("case 4M uniform NP6 q512 seed1", flush=True)
    base = K.Problem.create(4194304, "uniform", 512, np_=6, seed=1)
    ev = K.GpuEvaluator(base)
    pot1 = base.scatter(ev.evaluate()[0])
    err, _, _ = base.validate_gpu(ev, pot1, 1000)
    print(f"  sampled error {err:.3e}", flush=True)
    assert err < 1e-6
    del ev
    pts = base.points()
    rng = np.random.default_rng(5)
    w2 = rng.random(pts["den"].size)
    xyz = (pts["sx"], pts["sy"], pts["sz"])
    txyz = (pts["tx"], pts["ty"], pts["tz"])
    pots = []
    for w in (w2, pts["den"] - 2.0 * w2):
        p = K.Problem.from_points(*xyz, w, *txyz, pts_max=512, np_=6)
        e = K.GpuEvaluator(p)
        pots.append(p.scatter(e.evaluate()[0]))
        del e
    # phi(w1) = phi(w1 - 2 w2) + 2 phi(w2): every phase is linear in w.  Each
    # NP = 6 evaluation carries the pinv conditioning floor (oracle_lib.PARITY_RTOL,
    # ~2e-11 measured on the oracle itself); the combination adds three of them.
    r = rel(pots[1] + 2.0 * pots[0], pot1)
    print(f"  linearity {r:.3e}", flush=True)
    assert r < 3 * O.PARITY_RTOL[6]


def test_ellipsoid_1m_np4_graph():
    """1,048,576 points on the pole-clustered ellipsoid, seed 2, NP = 4,
    128 per box, FP64 captured in one HIP graph."""
    print("case 1M ellipsoid NP4 q128 seed2", flush=True)
    prob = K.Problem.create(1048576, "ellipsoid", 128, np_=4, seed=2)
    eager = K.GpuEvaluator(prob)
    graph = K.GpuEvaluator(prob, graph=True)
    a = eager.evaluate()[0]
    for _ in range(2):
        assert np.array_equal(graph.evaluate()[0], a)
    err, _, _ = prob.validate_gpu(eager, prob.scatter(a), 1000)
    print(f"  sampled error {err:.3e}", flush=True)
    assert err < 1e-4


@pytest.mark.parametrize("np_,cap", [(4, 1e-4), (3, 1e-3)])
def test_uniform_16m_mixed_vs_fp64(np_, cap):
    """16,777,216 uniform points, seed 3, 256 per box: mixed and FP64 on
    one tree, errors on one sample."""
    print(f"case 16M uniform NP{np_} q256 seed3", flush=True)
    prob = K.Problem.create(16777216, "uniform", 256, np_=np_, seed=3)
    fp = K.GpuEvaluator(prob)
    pf = prob.scatter(fp.evaluate()[0])
    ef, ids, direct = prob.validate_gpu(fp, pf, 1000)
    del fp
    mx = K.GpuEvaluator(prob, precision=K.PREC_MIXED)
    pm = prob.scatter(mx.evaluate()[0])
    em = rel(pm[ids], direct) * np.linalg.norm(direct) / np.linalg.norm(pm[ids])
    d = rel(pm, pf)
    print(f"  FP64 error {ef:.3e}, mixed error {em:.3e}, mixed vs FP64 {d:.3e}", flush=True)
    assert ef < cap and em < cap and abs(em - ef) <= 0.1 * ef and d < 1e-5
