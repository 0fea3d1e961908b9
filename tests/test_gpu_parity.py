"""GPU parity: the HIP FP64 path against the CPU oracle (exact kernel).

Both sides use the same tree (the product's and the oracle's trees are
bit-identical, see test_host_tree.py) and the same operator matrices (the
host-built UC2UE/UE2UC/DC2DE/DE2DC handed to kifmm_gpu_create, as the
reference's preproc uploads its host-built operators,
fmm/evaluate--cuda.cu:1597-1643,1777-1822).  Tolerance on the potentials:
oracle_lib.PARITY_RTOL (1e-12 relative L2 for NP <= 4; for NP = 6 three
times the measured summation-order floor of the oracle itself).
Intermediate equivalent densities sit before the far-field evaluation
that damps pinv-amplified rounding, so they get 1e3 times that bound.
Upward densities are compared from depth 2 on: the evaluator does not
form those of depth 0 and 1, which no phase reads.
"""
import numpy as np
import pytest

import kifmm_amd as K
import oracle_lib as O

pytestmark = pytest.mark.gpu

CASES = [
    # n, distribution, q, NP, seed, coincide
    (65536, "uniform", 64, 4, 0, 1),          # the reference's CPU case: source = target
    (32768, "uniform", 64, 6, 1, 0),
    (32768, "uniform", 32, 3, 5, 0),
    (65536, "ellipsoid", 32, 4, 2, 0),        # pole-clustered: W and X lists
    (32768, "ellipseUniformAngles", 16, 4, 2, 0),
    (20000, "gaussmix", 32, 4, 3, 0),
    (40000, "ellipsoid", 24, 6, 9, 1),
    (60000, "uniform", 300, 4, 4, 0),         # leaves of 130-300 targets: several U-list items per leaf
    (60000, "gaussmix", 256, 4, 3, 0),        # W and X lists over 1024 sources: chunked, k_wx_reduce
    (40000, "gaussmix", 512, 3, 1, 1),        # chunked X lists with coincident targets
]


def raw_ops(prob):
    o = prob.operators()
    ne, nuc = o.n_e, o.n_uc
    return [np.ctypeslib.as_array(p, shape=(k,)).copy() for p, k in
            ((o.uc2ue, ne * nuc), (o.ue2uc, 8 * nuc * ne), (o.dc2de, ne * ne), (o.de2dc, 8 * ne * ne))]


def run_pair(n, dist, q, np_, seed, c):
    print(f"case n={n} dist={dist} q={q} NP={np_} seed={seed} C={c}", flush=True)
    prob = K.Problem.create(n, dist, q, np_=np_, seed=seed, coincide=c)
    ev = K.GpuEvaluator(prob)
    pot_leaf, times = ev.evaluate()
    pot = prob.scatter(pot_leaf)
    orc = O.Oracle(n, dist, q, np_, seed, c)
    orc.set_operators(*raw_ops(prob))
    ref, _ = orc.evaluate(O.KERNEL_EXACT)
    return prob, ev, pot, orc, ref, times


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[1]}-n{c[0]}-q{c[2]}-np{c[3]}-C{c[5]}")
def test_potentials_match_oracle(case):
    prob, ev, pot, orc, ref, times = run_pair(*case)
    assert np.all(np.isfinite(pot))
    r = rel(pot, ref)
    tol = O.PARITY_RTOL[case[3]]
    print(f"  rel L2 {r:.3e} (bound {tol:.0e})  max rel {np.max(np.abs(pot - ref) / np.abs(ref)):.3e}  "
          f"total {times['total']:.3f} ms")
    # every phase's device vector against the oracle's
    uden, dchk, dden = orc.intermediates()
    tag = prob.tree()["tag"]
    src = (tag & K.KIFMM_SRCNODE) != 0
    trg = (tag & K.KIFMM_TRGNODE) != 0
    d = prob.tree()["depth"]
    for name, mine, theirs, mask in (("uden", ev.intermediate("uden"), uden, src & (d >= 2)),
                                     ("dchk", ev.intermediate("dchk"), dchk, trg & (d >= 2)),
                                     ("dden", ev.intermediate("dden"), dden, trg & (d >= 2))):
        if mask.sum() == 0:
            continue
        rr = rel(mine[mask], theirs[mask])
        print(f"  {name}: rel {rr:.3e}")
        assert rr <= 1e3 * tol, name
    assert r <= tol
    # FP64 U-list forms: Gram (close pairs listed, the default) and difference
    diff = K.GpuEvaluator(prob, inline_close=True)
    assert ev.work_counts()["close_pairs"] >= 0 and diff.work_counts()["close_pairs"] == -1
    rd = rel(prob.scatter(diff.evaluate()[0]), ref)
    print(f"  difference form: rel L2 {rd:.3e}")
    assert rd <= tol


@pytest.mark.parametrize("n,dist,q,np_,c", [(65536, "uniform", 64, 4, 1), (40000, "gaussmix", 128, 6, 0),
                                            (30000, "ellipsoid", 16, 3, 1), (50000, "uniform", 500, 4, 0)])
def test_gram_ulist_matches_difference_form(n, dist, q, np_, c):
    """The U-list alone, FP64: the Gram form (every pair closer than h/8
    listed and its term replaced by the difference form's) against the
    difference form, target by target.  A dropped or doubled close-pair
    term, a wrong scale or a wrong clamp shows at 1e-3 or more; measured:
    at most 5.5e-15 per target, 5e-16 relative L2."""
    print(f"case n={n} dist={dist} q={q} NP={np_} C={c}", flush=True)
    prob = K.Problem.create(n, dist, q, np_=np_, seed=3, coincide=c)
    ev = K.GpuEvaluator(prob)
    a, _ = ev.evaluate_phases(K.PHASE_U)
    b, _ = K.GpuEvaluator(prob, inline_close=True).evaluate_phases(K.PHASE_U)
    close = ev.work_counts()["close_pairs"]
    dev = np.abs(a - b) / np.abs(b)
    print(f"  close pairs {close:.0f}; max rel {dev.max():.3e}, rel L2 {rel(a, b):.3e}", flush=True)
    assert close > 0
    assert np.all(np.isfinite(a)) and dev.max() < 1e-13 and rel(a, b) < 1e-14


def test_sampled_error_vs_direct_sum():
    prob, ev, pot, orc, ref, _ = run_pair(65536, "uniform", 64, 4, 0, 1)
    err, ids, direct = prob.validate(pot, 1000)
    err_o, ids_o, _ = orc.validate(ref, 1000)
    assert np.array_equal(ids, ids_o)        # same generator stream continuation
    assert err < 1e-4 and abs(err - err_o) < 1e-9


def test_repeat_is_bitwise_deterministic():
    prob = K.Problem.create(32768, "ellipsoid", 32, np_=4, seed=4)
    ev = K.GpuEvaluator(prob)
    a, _ = ev.evaluate()
    b, _ = ev.evaluate()
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n,q", [(50, 64), (200, 8), (1000, 1)])
def test_tiny_and_ragged_trees(n, q):
    """root-only tree, a shallow tree, and q = 1 (one point per leaf)"""
    prob, ev, pot, orc, ref, _ = run_pair(n, "uniform", q, 4, 11, 0)
    assert rel(pot, ref) <= O.PARITY_RTOL[4]


def test_unsupported_np_is_an_error():
    prob = K.Problem.create(4096, "uniform", 64, np_=5, seed=1)
    with pytest.raises(K.KifmmError) as e:
        K.GpuEvaluator(prob)
    assert e.value.code == -4


def test_result_can_stay_on_device():
    prob = K.Problem.create(16384, "uniform", 64, np_=4, seed=2)
    ev = K.GpuEvaluator(prob)
    none, t = ev.evaluate(copy=False)
    assert none is None and t["copy"] < 0.5
    pot, _ = ev.evaluate(copy=True)
    assert np.array_equal(ev.intermediate("pot"), pot)


@pytest.mark.parametrize("prec", [K.PREC_FP64, K.PREC_MIXED])
def test_overlap_is_bitwise_identical(prec):
    """KIFMM_GPU_OVERLAP (the U-list on a second stream alongside up and V,
    joined before W, X and down) changes no sum: eager evaluations, graph
    replays and a U-list-only phase mask equal the one-stream results."""
    prob = K.Problem.create(60000, "gaussmix", 256, np_=4, seed=3)
    one = K.GpuEvaluator(prob, precision=prec)
    a = one.evaluate()[0]
    u = one.evaluate_phases(K.PHASE_U)[0]
    for graph in (False, True):
        ev = K.GpuEvaluator(prob, precision=prec, overlap=True, graph=graph)
        for _ in range(3):
            assert np.array_equal(ev.evaluate()[0], a)
        assert np.array_equal(ev.evaluate_phases(K.PHASE_U)[0], u)


def test_phase_subsets_after_a_full_evaluation_match_a_fresh_handle():
    """Only the buffers a call's phases do not assign are cleared (pot by the
    U-list, uden by the up pass, dden by the down pass; dchk always): after a
    full evaluation every phase subset must give the potentials and
    intermediates of a fresh handle, so no stale value leaks through."""
    prob = K.Problem.create(40000, "ellipsoid", 32, np_=4, seed=12)
    ev = K.GpuEvaluator(prob)
    ev.evaluate()
    for mask in (K.PHASE_U, K.PHASE_V | K.PHASE_DOWN, K.PHASE_UP | K.PHASE_V, K.PHASE_W | K.PHASE_X,
                 K.PHASE_UP | K.PHASE_U | K.PHASE_DOWN, K.PHASE_ALL):
        fresh = K.GpuEvaluator(prob)
        want = fresh.evaluate_phases(mask)[0]
        got = ev.evaluate_phases(mask)[0]
        assert np.array_equal(got, want), mask
        for name in ("uden", "dchk", "dden"):
            assert np.array_equal(ev.intermediate(name), fresh.intermediate(name)), (mask, name)
