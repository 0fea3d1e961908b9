"""Launch shapes and their autotuner (kifmm_gpu_set_tuning /
kifmm_gpu_autotune, the counterpart of the reference's BS / UPBS blocking
knobs, fmm/util.c:118-163).

Every U-list form (one block per item of <= 128 or <= 64 targets, or the
persistent per-XCD work queue) and every Hadamard frequency block (4, 8,
16) computes each output in the same order, so the potentials must be
bit-identical to the default shapes, eagerly and through a recaptured HIP
graph, with separate and coincident targets.  The autotuner must return
the fastest of the times it measured and leave that choice set."""
import numpy as np
import pytest

import kifmm_amd as K

pytestmark = pytest.mark.gpu

SHAPES = [(f, fb) for f in (K.P2P_STATIC, K.P2P_PERSISTENT, K.P2P_STATIC64) for fb in (4, 8, 16)]


@pytest.mark.parametrize("n,dist,q,np_,c", [(60000, "uniform", 64, 4, 0), (50000, "ellipsoid", 32, 6, 0),
                                             (30000, "gaussmix", 48, 3, 1)])
def test_every_shape_gives_identical_potentials(n, dist, q, np_, c):
    print(f"case n={n} dist={dist} q={q} NP={np_} C={c}", flush=True)
    prob = K.Problem.create(n, dist, q, np_=np_, seed=4, coincide=c)
    ev = K.GpuEvaluator(prob)
    ref, _ = ev.evaluate()
    gr = K.GpuEvaluator(prob, graph=True)
    for f, fb in SHAPES:
        ev.set_tuning(f, fb)
        assert ev.tuning() == {"p2p_form": f, "hadamard_fb": fb}
        got, _ = ev.evaluate()
        assert np.array_equal(got, ref), (f, fb)
        gr.set_tuning(f, fb)
        for _ in range(2):
            got, _ = gr.evaluate()
            assert np.array_equal(got, ref), ("graph", f, fb)


def test_mixed_path_hadamard_blocks_and_forms():
    prob = K.Problem.create(40000, "uniform", 64, np_=4, seed=5)
    ev = K.GpuEvaluator(prob, precision=K.PREC_MIXED)
    ref, _ = ev.evaluate()
    for fb in (4, 16):
        ev.set_tuning(K.P2P_STATIC, fb)
        got, _ = ev.evaluate()
        assert np.array_equal(got, ref), fb
    with pytest.raises(K.KifmmError) as e:
        ev.set_tuning(K.P2P_PERSISTENT, 8)
    assert e.value.code == -4


def test_invalid_tuning_is_rejected():
    ev = K.GpuEvaluator(K.Problem.create(5000, "uniform", 64, np_=4, seed=6))
    for f, fb in ((3, 8), (-1, 8), (0, 5), (0, 0)):
        with pytest.raises(K.KifmmError) as e:
            ev.set_tuning(f, fb)
        assert e.value.code == -1
    assert ev.tuning() == {"p2p_form": K.P2P_STATIC, "hadamard_fb": 8}


@pytest.mark.parametrize("precision", [K.PREC_FP64, K.PREC_MIXED])
def test_autotune_keeps_the_fastest(precision):
    prob = K.Problem.create(200000, "ellipsoid", 64, np_=4, seed=7)
    ev = K.GpuEvaluator(prob, precision=precision)
    ref, _ = ev.evaluate()
    best, ms = ev.autotune(reps=2)
    print(f"  times {np.round(ms, 4)} -> {best}", flush=True)
    nforms = 3 if precision == K.PREC_FP64 else 1
    assert np.all(ms[:nforms] > 0) and np.all(ms[nforms:3] == -1) and np.all(ms[3:] > 0)
    assert best["p2p_form"] == int(np.argmin(ms[:nforms]))
    assert best["hadamard_fb"] == (4, 8, 16)[int(np.argmin(ms[3:]))]
    assert ev.tuning() == best
    got, _ = ev.evaluate()
    assert np.array_equal(got, ref)
