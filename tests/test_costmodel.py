"""Cost model, host side: the numpy network restatement is pinned by finite
differences (so it can check the HIP gradients), features are finite and
deterministic, and the committed weights file loads without executing
anything (JSON)."""
import numpy as np

import costmodel as CM
import kifmm_amd as K
import mlp_ref as R


def test_numpy_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = CM.init_params(1).astype(np.float64)
    X = rng.standard_normal((20, 16))
    T = rng.standard_normal((20, 4))
    _, g = R.loss_and_grad(p, X, T)
    idx = rng.choice(p.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (R.loss_and_grad(p + e, X, T)[0] - R.loss_and_grad(p - e, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), i


def test_parameter_count_and_init():
    p = CM.init_params(0)
    assert p.size == CM.NPARAMS == 16 * 64 + 64 + 64 * 64 + 64 + 64 * 4 + 4
    assert np.all(p[16 * 64:16 * 64 + 64] == 0)


def test_features_are_finite_and_deterministic():
    p = K.Problem.create(30000, "ellipsoid", 32, np_=4, seed=3)
    a = CM.features(p, 32, K.PREC_FP64)
    b = CM.features(p, 32, K.PREC_FP64)
    c = CM.features(p, 32, K.PREC_MIXED)
    assert a.shape == (16,) and np.all(np.isfinite(a)) and np.array_equal(a, b)
    assert c[12] == 1.0 and a[12] == 0.0 and a[11] == 4.0 and a[13] == 5.0


def test_saved_model_roundtrip(tmp_path):
    m = CM.CostModel(CM.init_params(2), np.zeros(16), np.ones(16), np.zeros(4), np.ones(4), {"rows": 1})
    f = tmp_path / "w.json"
    m.save(f)
    m2 = CM.CostModel.load(f)
    assert np.array_equal(m2.params, m.params) and m2.meta["rows"] == 1


def test_committed_report_meets_its_claims():
    """The held-out report of the committed weights (train_costmodel.py):
    16 inputs not in the training set, each with the model's pick, the best
    of the exhaustive sweep and the fixed q = 64 FP64 default, timed, and
    each pick's sampled error against the FP64 direct sum.  The pick is near
    the best (geomean within 5 %, never worse than 1.25x), gains over the
    default, and every sampled error is at the FMM's own level for its NP."""
    import json
    from pathlib import Path
    rep = json.loads((Path(CM.__file__).parent / "costmodel_report.json").read_text())
    held = rep["held_out"]
    assert len(held) == 16
    bound = {3: 5e-4, 4: 1e-4, 6: 1e-6}
    for r in held:
        for k in ("chosen", "best", "verify2"):
            e = r[k]["sampled_rel_err"]
            assert e is not None and 0 < e <= bound[r["np"]], (r["n"], r["dist"], k, e)
        assert r["chosen"]["ms"] >= r["best"]["ms"] and r["chosen_over_best"] <= 1.25
    g = lambda k: float(np.exp(np.mean(np.log([r[k] for r in held]))))
    assert g("chosen_over_best") <= 1.05 and g("verify2_over_best") <= 1.01
    assert g("gain_over_default") >= 1.5
    assert rep["model"]["rows"] > 500 and rep["model"]["final_loss"] < rep["model"]["first_loss"]
