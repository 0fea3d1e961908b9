"""GPU sampled-direct validator against the host one (fmm/driver.cc:141-206):
identical sample ids (the generator stream continued the same way), direct
sums equal to 1e-13 relative (compensated FP64 against plain FP64 over a
different order), identical error to 1e-9 relative; r = 0 excluded for
coincident points."""
import numpy as np
import pytest

import kifmm_amd as K

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,dist,c", [(200000, "uniform", 0), (100000, "gaussmix", 1), (50000, "ellipsoid", 0)])
def test_gpu_validate_matches_host(n, dist, c):
    print(f"case n={n} dist={dist} C={c}", flush=True)
    a = K.Problem.create(n, dist, 64, np_=4, seed=21, coincide=c)
    b = K.Problem.create(n, dist, 64, np_=4, seed=21, coincide=c)
    ev = K.GpuEvaluator(b)
    pot = b.scatter(ev.evaluate()[0])
    e_host, ids_h, d_host = a.validate(pot, 1000)
    e_gpu, ids_g, d_gpu = b.validate_gpu(ev, pot, 1000)
    assert np.array_equal(ids_h, ids_g)
    assert np.max(np.abs(d_gpu - d_host) / np.abs(d_host)) < 1e-13
    assert abs(e_gpu - e_host) <= 1e-9 * e_host


def test_gpu_direct_known_answer():
    p = K.Problem.create(1000, "uniform", 64, np_=4, seed=3)
    ev = K.GpuEvaluator(p)
    pts = p.points()
    t = np.array([3.0]), np.array([0.25]), np.array([-2.0])
    out = np.zeros(1)
    dp = lambda a: a.ctypes.data_as(K.C.POINTER(K.C.c_double))
    K._check(K.lib.kifmm_gpu_direct(ev._h, 1, dp(t[0]), dp(t[1]), dp(t[2]), dp(out)))
    r = np.sqrt((t[0] - pts["sx"]) ** 2 + (t[1] - pts["sy"]) ** 2 + (t[2] - pts["sz"]) ** 2)
    assert out[0] == pytest.approx((pts["den"] / r).sum() / (4 * np.pi), rel=1e-14)
