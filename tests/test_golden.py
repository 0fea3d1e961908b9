"""Committed fixtures (tests/golden/, written by tests/golden/make_golden.py)."""
from pathlib import Path

import ctypes as C
import numpy as np

import kifmm_amd as K
import oracle_lib as O

G = Path(__file__).resolve().parent / "golden"


def test_drand48_fixture():
    f = np.load(G / "drand48_glibc.npz", allow_pickle=False)
    for seed in (0, 1, 2, 3, 26):
        assert np.array_equal(O.drand48(seed, 64), f[f"seed{seed}"])
        out = np.zeros(64)
        K.lib.kifmm_drand48(seed, 64, out.ctypes.data_as(C.POINTER(C.c_double)))
        assert np.array_equal(out, f[f"seed{seed}"])


def test_pinv_fixture_np3():
    f = np.load(G / "pinv_np3_numpy.npz", allow_pickle=False)
    p = K.Problem.create(500, "uniform", 64, np_=3, seed=1).operator_arrays()
    assert np.linalg.norm(p["uc2ue"] - f["uc2ue"]) / np.linalg.norm(f["uc2ue"]) < 1e-12
    assert np.linalg.norm(p["dc2de"] - f["dc2de"]) / np.linalg.norm(f["dc2de"]) < 1e-12
    o = O.Oracle(500, "uniform", 64, 3, 1, 0)
    o.build_operators(O.KERNEL_EXACT)
    uc2ue = o.get_operators()[0].reshape(K.n_uc(3), K.n_e(3)).T
    assert np.linalg.norm(uc2ue - f["uc2ue"]) / np.linalg.norm(f["uc2ue"]) < 1e-12


def test_rfftn_fixture():
    f = np.load(G / "rfftn_N6_numpy.npz", allow_pickle=False)
    assert np.max(np.abs(O.rfft3(f["grid"]) - f["spectrum"])) < 1e-13


def test_oracle_regression_fixture():
    f = np.load(G / "oracle_small.npz", allow_pickle=False)
    o = O.Oracle(2000, "uniform", 16, 4, 0, 1)
    o.build_operators(O.KERNEL_EXACT)
    pot, _ = o.evaluate(O.KERNEL_EXACT)
    assert np.linalg.norm(pot - f["pot"]) / np.linalg.norm(f["pot"]) < 1e-13
    err, ids, _ = o.validate(pot, 100)
    assert np.array_equal(ids, f["ids"]) and abs(err - float(f["err"])) < 1e-12
