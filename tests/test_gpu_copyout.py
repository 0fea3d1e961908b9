"""Copy-out, density updates and the create-time close-pair list.

The potentials come back to the host as part of every evaluation (the
reference's copy_trg_val__gpu, fmm/cuda.cu:2543-2565, inside run(),
fmm/run--cuda.cu:319-324).  The U-list runs last, in chunks whose copies
overlap the later chunks; whatever the route (pinned buffer, caller buffer,
graph replay, either stream layout, any U-list form) the host must receive
bit for bit what the device holds.  New densities on an existing handle must
give bit for bit what a fresh handle built with them gives.  The close-pair
list (U-list pairs closer than 1/8 of the target leaf's half-width, found
at create time by the near-box search) must hold, per target and in U-list
order, exactly the pairs a brute force over the tree's own lists finds."""
import numpy as np
import pytest

import kifmm_amd as K

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [K.PREC_FP64, K.PREC_MIXED], ids=["fp64", "mixed"])
def test_copy_out_is_the_device_result(prec):
    n = 300000  # above the size at which the U-list is split into copy-out chunks
    print(f"case uniform n={n} q=64 NP=4 prec={prec}", flush=True)
    prob = K.Problem.create(n, "uniform", 64, np_=4, seed=11)
    ev = K.GpuEvaluator(prob, precision=prec)
    assert ev.work_counts()["copy_chunks"] > 1
    ev.evaluate(copy=False)
    dev = ev.intermediate("pot")
    view, t = ev.evaluate(copy="view")
    assert np.array_equal(view, dev)
    assert t["total"] > 0 and t["copy"] >= 0
    own = np.full(ev.n_trg, np.nan)
    ev.evaluate_into(own)
    assert np.array_equal(own, dev)
    got, _ = ev.evaluate()  # numpy copy of the pinned buffer
    assert np.array_equal(got, dev)
    for ov in (1, 2, 3, 4, 5, 6):  # on the second stream from the start: all, first, first two; 4 - 6: one stream
        ev.set_tuning(K.P2P_STATIC, 8, K.HAD_MFMA, ov)
        got, t = ev.evaluate()
        assert np.array_equal(got, dev), ov
        own[:] = np.nan
        ev.evaluate_into(own)
        assert np.array_equal(own, dev), ov
        ev.evaluate(copy=False)
        assert np.array_equal(ev.intermediate("pot"), dev), ov
    forms = (K.P2P_PERSISTENT, K.P2P_STATIC64) if prec == K.PREC_FP64 else (K.P2P_SPLIT,)
    split = None
    for f in forms:
        for ov in (0, 1, 2, 4):
            ev.set_tuning(f, 8, K.HAD_MFMA, ov)
            got, _ = ev.evaluate()
            if f == K.P2P_SPLIT:  # its own sums (rounding-level differences), the same in every layout
                split = got if split is None else split
                assert np.array_equal(got, split), (f, ov)
            else:
                assert np.array_equal(got, dev), (f, ov)
    gr = K.GpuEvaluator(prob, precision=prec, graph=True)
    for k in range(7):
        gr.set_tuning(K.P2P_STATIC, 8, K.HAD_MFMA, k)  # a recapture per layout
        got, _ = gr.evaluate()
        assert np.array_equal(got, dev), ("graph replay", k)
        gr.evaluate(copy=False)
        assert np.array_equal(gr.intermediate("pot"), dev)
        got, _ = gr.evaluate()
        assert np.array_equal(got, dev), ("graph replay", k)
        ev.evaluate()


def test_phase_subsets_copy_out():
    """A subset of phases copies back what it computed (U-list last or absent)."""
    prob = K.Problem.create(280000, "ellipsoid", 64, np_=4, seed=12)
    ev = K.GpuEvaluator(prob)
    for mask in (K.PHASE_U, K.PHASE_ALL & ~K.PHASE_U, K.PHASE_UP | K.PHASE_V | K.PHASE_DOWN, K.PHASE_ALL):
        ev.evaluate_phases(mask, copy=False)
        dev = ev.intermediate("pot")
        got, _ = ev.evaluate_phases(mask)
        assert np.array_equal(got, dev), mask
    u, _ = ev.evaluate_phases(K.PHASE_U)
    far, _ = ev.evaluate_phases(K.PHASE_ALL & ~K.PHASE_U)
    full, _ = ev.evaluate_phases(K.PHASE_ALL)
    assert np.array_equal(full, u + far)  # the U-list's write is u + pot


@pytest.mark.parametrize("prec,dist,c", [(K.PREC_FP64, "ellipsoid", 1), (K.PREC_MIXED, "uniform", 0),
                                         (K.PREC_FP64, "gaussmix", 0)])
def test_set_densities_matches_a_fresh_handle(prec, dist, c):
    n, q, np_ = 60000, 48, 4
    print(f"case {dist} n={n} q={q} NP={np_} C={c} prec={prec}", flush=True)
    a = K.Problem.create(n, dist, q, np_=np_, seed=13, coincide=c)
    pts = a.points()
    rng = np.random.default_rng(5)
    den2 = rng.uniform(-1.0, 1.0, n)
    if c:
        b = K.Problem.from_points(pts["sx"], pts["sy"], pts["sz"], den2, pts_max=q, np_=np_)
    else:
        b = K.Problem.from_points(pts["sx"], pts["sy"], pts["sz"], den2, pts["tx"], pts["ty"], pts["tz"],
                                  pts_max=q, np_=np_)
    ta, tb = a.tree(), b.tree()
    assert np.array_equal(ta["sx"], tb["sx"]) and np.array_equal(ta["src_beg"], tb["src_beg"])
    ev = K.GpuEvaluator(a, precision=prec)
    first, _ = ev.evaluate()
    ev.set_densities(tb["sden"])
    got, _ = ev.evaluate()
    want, _ = K.GpuEvaluator(b, precision=prec).evaluate()
    assert np.array_equal(got, want)
    ev.set_densities(ta["sden"])
    again, _ = ev.evaluate()
    assert np.array_equal(again, first)


@pytest.mark.parametrize("n,dist,q,c", [(6000, "uniform", 32, 0), (5000, "ellipsoid", 16, 1),
                                         (4000, "gaussmix", 8, 0), (3000, "uniform", 1, 1)])
def test_close_list_is_the_brute_force_list(n, dist, q, c):
    """Per target, the listed sources are the U-list sources at scaled
    distance r' < 1 in U-list order; only pairs within 1e-9 of the threshold
    may go either way (numpy's rounding differs from the kernel's fma)."""
    prob = K.Problem.create(n, dist, q, np_=4, seed=14, coincide=c)
    ev = K.GpuEvaluator(prob)
    lst = ev.close_list()
    assert lst is not None
    ptr, idx = lst
    t = prob.tree()
    S = np.stack([t["sx"], t["sy"], t["sz"]], 1)
    T = np.stack([t["tx"], t["ty"], t["tz"]], 1)
    cen = np.stack([t["path"][:, k] for k in range(3)], 1)
    assert ptr[0] == 0 and ptr[-1] == idx.size and np.all(np.diff(ptr) >= 0)
    checked = 0
    for i in range(t["n_nodes"]):
        if t["child"][i] != -1 or t["trg_num"][i] == 0:
            continue
        h = 2.0 ** -int(t["depth"][i])
        g = -1.0 + 2.0 * (cen[i] + 0.5) * h
        sc = 8.0 / h
        srcs = np.concatenate([np.arange(t["src_beg"][s], t["src_beg"][s] + t["src_num"][s])
                               for s in t["u_idx"][t["u_ptr"][i]:t["u_ptr"][i + 1]]] + [np.zeros(0, np.int64)])
        sp = (S[srcs] - g) * sc
        for k in range(t["trg_beg"][i], t["trg_beg"][i] + t["trg_num"][i]):
            d = (T[k] - g) * sc - sp
            r2 = np.sum(d * d, axis=1)
            sure, maybe = srcs[r2 < 1.0 - 1e-9], srcs[r2 < 1.0 + 1e-9]
            got = idx[ptr[k]:ptr[k + 1]]
            assert np.all(np.isin(sure, got)) and np.all(np.isin(got, maybe)), k
            assert np.array_equal(got, maybe[np.isin(maybe, got)]), k  # U-list order
            checked += got.size
    assert checked == idx.size and checked > 0
    ms = ev.create_times()
    print(f"  {checked} close pairs; create ms {ms}", flush=True)
    assert all(v >= 0 for v in ms.values()) and ms["total"] >= ms["close_pairs"]
