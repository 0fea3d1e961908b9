"""Randomized parity sweep: 24 small inputs drawn with a fixed seed across
every generator, points per box from 1 to 200, NP 3 / 4 / 6, separate and
coincident targets, each evaluated under a randomly drawn launch shape
(U-list form, Hadamard form and frequency block) and stream layout
(one stream, or the U-list overlapped on a second one).  The FP64 potentials
must meet the parity bar against the oracle (oracle_lib.PARITY_RTOL); for
NP <= 4 the mixed path must be within 1e-5 of the oracle and finite.  The
fixed cases of test_gpu_parity.py pin specific shapes; this one looks for
combinations nobody wrote down."""
import numpy as np
import pytest

import kifmm_amd as K
import oracle_lib as O

pytestmark = pytest.mark.gpu

DISTS = ("uniform", "ellipsoid", "ellipseUniformAngles", "gaussmix")


def draw_cases(count=24, seed=2026):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        q = int(rng.choice([1, 3, 8, 17, 64, 200]))
        nmax = 3000 if q <= 3 else 30000
        n = int(np.exp(rng.uniform(np.log(300), np.log(nmax))))
        out.append((n, DISTS[i % 4], q, int(rng.choice([3, 4, 6])), int(rng.integers(0, 1000)),
                    int(rng.integers(0, 2)), int(rng.integers(0, 3)), int(rng.choice([4, 8, 16])),
                    int(rng.integers(0, 2)), bool(rng.integers(0, 2))))
    return out


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("case", draw_cases(), ids=lambda c: f"{c[1]}-n{c[0]}-q{c[2]}-np{c[3]}-C{c[5]}")
def test_random_inputs_and_shapes_match_oracle(case):
    n, dist, q, np_, seed, c, form, fb, had, ovl = case
    print(f"case n={n} dist={dist} q={q} NP={np_} seed={seed} C={c} p2p_form={form} fb={fb} had={had} "
          f"overlap={ovl}", flush=True)
    prob = K.Problem.create(n, dist, q, np_=np_, seed=seed, coincide=c)
    o = prob.operators()
    orc = O.Oracle(n, dist, q, np_, seed, c)
    orc.set_operators(*[np.ctypeslib.as_array(p, shape=(k,)).copy() for p, k in
                        ((o.uc2ue, o.n_e * o.n_uc), (o.ue2uc, 8 * o.n_uc * o.n_e), (o.dc2de, o.n_e ** 2),
                         (o.de2dc, 8 * o.n_e ** 2))])
    ref, _ = orc.evaluate(O.KERNEL_EXACT)
    ev = K.GpuEvaluator(prob, overlap=ovl)
    ev.set_tuning(form, fb, had)
    got = prob.scatter(ev.evaluate()[0])
    r = rel(got, ref)
    print(f"  FP64 vs oracle {r:.2e}", flush=True)
    assert r < O.PARITY_RTOL[np_]
    if np_ <= 4:
        mx = K.GpuEvaluator(prob, precision=K.PREC_MIXED, overlap=ovl)
        mx.set_tuning(K.P2P_STATIC, fb, had)
        pm = prob.scatter(mx.evaluate()[0])
        rm = rel(pm, ref)
        print(f"  mixed vs oracle {rm:.2e}", flush=True)
        assert np.all(np.isfinite(pm)) and rm < 1e-5
