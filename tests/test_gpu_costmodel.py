"""Cost model on the GPU: the HIP forward, backward and Adam kernels
against the numpy restatement (tests/mlp_ref.py, itself pinned by finite
differences), and what the trained network gains over running without it."""
import numpy as np
import pytest

import costmodel as CM
import kifmm_amd as K
import mlp_ref as R

pytestmark = pytest.mark.gpu


def data(n=96, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, 16)).astype(np.float32)
    T = np.tanh(X[:, :4] * 0.7 + 0.3 * X[:, 4:8]).astype(np.float32)
    return X, T


def test_forward_and_gradient_match_numpy():
    X, T = data()
    p = CM.init_params(3)
    net = CM.Network(p)
    Y = net.predict(X)
    Yr, _ = R.forward(p.astype(np.float64), X.astype(np.float64))
    assert np.max(np.abs(Y - Yr)) <= 1e-4 * max(1.0, np.max(np.abs(Yr)))
    g, loss = net.gradient(X, T)
    lr, gr = R.loss_and_grad(p.astype(np.float64), X.astype(np.float64), T.astype(np.float64))
    assert abs(loss - lr) <= 1e-5 * max(1.0, lr)
    assert np.linalg.norm(g - gr) <= 1e-4 * np.linalg.norm(gr)
    g2, _ = net.gradient(X, T)
    assert np.array_equal(g, g2)          # fixed reduction order


def test_adam_steps_match_numpy():
    X, T = data(64, 1)
    p = CM.init_params(4)
    net = CM.Network(p)
    losses = net.train(X, T, 25, 1e-3)
    pr, lr = R.adam(p, X.astype(np.float64), T.astype(np.float64), 25, 1e-3)
    assert np.allclose(losses, lr, rtol=1e-3, atol=1e-6)
    assert np.linalg.norm(net.params() - pr) <= 1e-3 * np.linalg.norm(pr)
    assert losses[-1] < losses[0]


def test_trained_model_gains_over_default_q():
    """Held-out inputs (not among the training seeds): the q and path the
    committed model picks against the reference's default practice (a fixed
    q = 64 on the FP64 path) and against the best of an exhaustive sweep."""
    model = CM.CostModel.load()
    if model is None:
        pytest.skip("no trained weights committed")
    tuner = CM.Tuner(model)
    rows = []
    for n, dist, np_, seed in ((1 << 19, "uniform", 4, 9001), (1 << 19, "ellipsoid", 4, 9002),
                               (1 << 18, "gaussmix", 6, 9003)):
        base = K.Problem.create(n, dist, 64, np_=np_, seed=seed)
        pts = base.points()
        pts = (pts["sx"], pts["sy"], pts["sz"], pts["den"], pts["tx"], pts["ty"], pts["tz"])
        q, prec, _, how = tuner.choose(pts, np_)
        assert how == "model"
        chosen = K.Problem.from_points(*pts, pts_max=q, np_=np_)
        _, t_model = CM.phase_times(chosen, prec, steps=5)
        _, t_default = CM.phase_times(base, K.PREC_FP64, steps=5)
        print(f"{dist} n={n} NP={np_}: model q={q} prec={prec} {t_model:.3f} ms; default q=64 FP64 {t_default:.3f} ms")
        rows.append(t_default / t_model)
    # the committed report's held-out geomean is 2.08 (16 inputs, least 1.42);
    # 1.5 leaves room for box-to-box timing spread on these three
    assert np.exp(np.mean(np.log(rows))) >= 1.5


def test_verified_choice_is_one_of_the_best_ranked():
    """Tuner(verify=2): the model ranks every (q, path), the two best-ranked
    are timed, and the faster of them is returned (measured, not predicted)."""
    model = CM.CostModel.load()
    if model is None:
        pytest.skip("no trained weights committed")
    n, dist, np_, seed = 1 << 18, "uniform", 4, 9004
    base = K.Problem.create(n, dist, 64, np_=np_, seed=seed)
    pts = base.points()
    pts = (pts["sx"], pts["sy"], pts["sz"], pts["den"], pts["tx"], pts["ty"], pts["tz"])
    tuner = CM.Tuner(model)
    rows, keys = [], []
    ops = base.operators()
    for q in tuner.qs:
        prob = K.Problem.from_points(*pts, pts_max=q, np_=np_)
        for pr in CM.paths_for(np_):
            rows.append(CM.features(prob, q, pr, ops))
            keys.append((q, pr))
    order = np.argsort(model.predict_ms(np.stack(rows)).sum(1), kind="stable")
    top2 = {keys[int(i)] for i in order[:2]}
    q, prec, ms, how = tuner.choose(pts, np_, verify=2)
    print(f"top-2 {sorted(top2)} -> q={q} prec={prec} {ms:.3f} ms", flush=True)
    assert how == "model+verify" and (q, prec) in top2 and ms > 0
