"""kifmm_gpu_direct_noalloc (no counterpart in the reference: the direct
FP64 sums of kifmm_gpu_direct, for callers that cannot allocate, with its
staging in the library's file-scope device scratch) against kifmm_gpu_direct
on the same handle: two calls in a row, the second with fewer points, bit for
bit; points beyond the scratch's capacity are an error, not a fault."""
import numpy as np
import pytest

import kifmm_amd as K

pytestmark = pytest.mark.gpu

CAP = 16384  # KIFMM_DIRECT_NOALLOC_MAX


def test_noalloc_direct_matches_direct():
    prob = K.Problem.create(30000, "ellipsoid", 32, np_=4, seed=21, coincide=1)
    ev = K.GpuEvaluator(prob)
    rng = np.random.default_rng(3)
    for nt in (5000, 1200):  # the second call smaller: stale scratch beyond nt must not matter
        print(f"case noalloc direct nt={nt}", flush=True)
        t = rng.uniform(-1.0, 1.0, (3, nt))
        t[:, :10] = np.stack([prob.tree()[k][:10] for k in ("sx", "sy", "sz")])  # r = 0 pairs too
        want = ev.direct(*t)
        got = ev.direct(*t, noalloc=True)
        assert np.array_equal(got, want)
        assert np.all(np.isfinite(got))
    t = rng.uniform(-1.0, 1.0, (3, CAP))
    assert np.array_equal(ev.direct(*t, noalloc=True), ev.direct(*t))
    with pytest.raises(K.KifmmError) as e:
        ev.direct(*rng.uniform(-1.0, 1.0, (3, CAP + 1)), noalloc=True)
    assert e.value.code == -1 and "scratch" in str(e.value)


def test_noalloc_direct_on_a_second_handle():
    """Another handle on the same device uses the same scratch (its own
    sources): results equal its own allocating call."""
    a = K.GpuEvaluator(K.Problem.create(8000, "uniform", 64, np_=4, seed=22))
    b = K.GpuEvaluator(K.Problem.create(9000, "gaussmix", 64, np_=4, seed=23))
    t = np.random.default_rng(4).uniform(-1.0, 1.0, (3, 3000))
    ra, rb = a.direct(*t, noalloc=True), b.direct(*t, noalloc=True)
    assert np.array_equal(ra, a.direct(*t)) and np.array_equal(rb, b.direct(*t))
    assert not np.array_equal(ra, rb)
