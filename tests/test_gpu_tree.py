"""GPU octree / interaction-list build against the host build (which follows
fmm/input.cc:203-460 and is itself checked against the oracle and the list
definitions in test_host_tree.py): every node array, every U/V/W/X list and
the gathered (permuted) points must be identical, bit for bit, across the
generators, separate and coincident targets, dyadic-grid points that sit
exactly on box centres and faces, ragged and tiny inputs, and the full-size
4M-point workload.  A problem built on the GPU must also evaluate to the same
potential as one built on the host."""
import time

import numpy as np
import pytest

import kifmm_amd as K

pytestmark = pytest.mark.gpu


def assert_same_tree(a, b):
    ta, tb = a.tree(), b.tree()
    assert ta.keys() == tb.keys()
    for k in ta:
        va, vb = ta[k], tb[k]
        if isinstance(va, np.ndarray):
            assert va.shape == vb.shape, k
            assert np.array_equal(va, vb), k
        else:
            assert va == vb, k


@pytest.mark.parametrize("n,dist,q,np_,c", [
    (20000, "uniform", 64, 4, 1),
    (20000, "uniform", 64, 4, 0),
    (50000, "ellipsoid", 32, 6, 0),
    (30000, "gaussmix", 16, 3, 0),
    (30000, "gaussmix", 64, 4, 1),
    (30000, "ellipseUniformAngles", 64, 4, 0),
    (3000, "uniform", 1, 3, 0),
    (1, "uniform", 64, 4, 0),
    (64, "uniform", 64, 4, 0),
    (65, "uniform", 64, 4, 0),
])
def test_gpu_tree_matches_host(n, dist, q, np_, c):
    print(f"case n={n} dist={dist} q={q} NP={np_} C={c}", flush=True)
    a = K.Problem.create(n, dist, q, np_=np_, seed=5, coincide=c)
    b = K.Problem.create(n, dist, q, np_=np_, seed=5, coincide=c, device=0)
    assert_same_tree(a, b)


def test_gpu_tree_dyadic_grid_points():
    """Points on the dyadic grid -1 + k/4 (box faces and centres at every
    level up to 3), including the faces x = +-1, once each, and targets
    shifted by a quarter box: the >= centre rule must be applied exactly."""
    g = -1.0 + np.arange(9) / 4.0
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    sx, sy, sz = X.ravel(), Y.ravel(), Z.ravel()
    rng = np.random.default_rng(3)
    den = rng.uniform(-1, 1, sx.size)
    tx, ty, tz = (np.clip(v + 0.0625, -1, 1) for v in (sx, sy, sz))
    for q in (1, 3, 8):
        print(f"case dyadic grid q={q}", flush=True)
        a = K.Problem.from_points(sx, sy, sz, den, tx, ty, tz, pts_max=q, np_=4)
        b = K.Problem.from_points(sx, sy, sz, den, tx, ty, tz, pts_max=q, np_=4, device=0)
        assert_same_tree(a, b)


def test_gpu_tree_duplicates_within_q():
    rng = np.random.default_rng(4)
    pts = rng.uniform(-1, 1, (3, 2000))
    pts[:, :40] = pts[:, :1]  # 40 coincident sources, q = 40 keeps them in one leaf
    den = rng.uniform(0, 1, 2000)
    a = K.Problem.from_points(*pts, den, pts_max=40, np_=4)
    b = K.Problem.from_points(*pts, den, pts_max=40, np_=4, device=0)
    assert_same_tree(a, b)


def test_gpu_tree_too_deep_fails_like_host():
    pts = np.zeros((3, 50)) + 0.3
    den = np.ones(50)
    with pytest.raises(K.KifmmError, match="30 levels"):
        K.Problem.from_points(*pts, den, pts_max=8, np_=4, device=0)
    with pytest.raises(K.KifmmError, match="30 levels"):
        K.Problem.from_points(*pts, den, pts_max=8, np_=4)


def test_gpu_built_problem_evaluates_identically():
    a = K.Problem.create(100000, "ellipsoid", 64, np_=4, seed=8)
    b = K.Problem.create(100000, "ellipsoid", 64, np_=4, seed=8, device=0)
    pa = K.GpuEvaluator(a).evaluate()[0]
    pb = K.GpuEvaluator(b).evaluate()[0]
    assert np.array_equal(pa, pb)


@pytest.mark.parametrize("n,dist,q,np_", [(1 << 22, "uniform", 512, 6), (1 << 20, "ellipsoid", 64, 4)])
def test_gpu_tree_full_size(n, dist, q, np_):
    print(f"case n={n} dist={dist} q={q}", flush=True)
    t0 = time.perf_counter()
    a = K.Problem.create(n, dist, q, np_=np_, seed=1)
    t1 = time.perf_counter()
    b = K.Problem.create(n, dist, q, np_=np_, seed=1, device=0)
    t2 = time.perf_counter()
    sa, sb = a.setup_times(), b.setup_times()
    print(f"  host tree+lists {sa['tree_lists'] * 1e3:.1f} ms, GPU {sb['tree_lists'] * 1e3:.1f} ms "
          f"(create {t1 - t0:.2f} s / {t2 - t1:.2f} s)", flush=True)
    assert_same_tree(a, b)
