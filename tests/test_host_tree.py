"""Host side: generators, octree, interaction lists and gather.

The product's tree and lists must be identical to the reference's
(fmm/input.cc:203-460): checked bit-for-bit against the oracle's
restatement, and independently against the textbook list definitions by
brute force on small non-uniform trees.
"""
import numpy as np
import pytest

import kifmm_amd as K
import oracle_lib as O

KEYS = ("parent", "child", "depth", "tag", "src_beg", "src_num", "trg_beg", "trg_num", "path", "level_start",
        "u_ptr", "u_idx", "v_ptr", "v_idx", "w_ptr", "w_idx", "x_ptr", "x_idx")

CASES = [
    (65536, "uniform", 64, 4, 0, 1),
    (40000, "uniform", 64, 6, 1, 0),
    (30000, "ellipseUniformAngles", 16, 4, 2, 0),
    (50000, "ellipsoid", 32, 3, 2, 0),
    (20000, "gaussmix", 24, 4, 3, 0),
    (5000, "ellipsoid", 2, 4, 5, 1),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[1]}-{c[0]}-q{c[2]}-C{c[5]}")
def test_tree_and_lists_match_oracle(case):
    n, dist, q, np_, seed, c = case
    p = K.Problem.create(n, dist, q, np_=np_, seed=seed, coincide=c)
    o = O.Oracle(n, dist, q, np_, seed, c)
    a, b = p.tree(), o.tree()
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k
    pa, pb = p.points(), o.points()
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k
    sp, tp = o.perms()
    assert np.array_equal(a["sx"], pa["sx"][sp]) and np.array_equal(a["sden"], pa["den"][sp])
    assert np.array_equal(a["tz"], pa["tz"][tp])


def test_tree_invariants():
    q = 32
    p = K.Problem.create(30000, "ellipsoid", q, np_=4, seed=8)
    t = p.tree()
    nn = t["n_nodes"]
    leaf = t["child"] == -1
    assert t["src_num"][leaf].max() <= q and t["src_num"].sum() == 30000 and t["trg_num"].sum() == 30000
    inner = np.where(~leaf)[0]
    assert np.all((t["child"][inner] - 1) % 8 == 0)                         # children contiguous, in blocks
    for k in inner:
        ch = t["child"][k] + np.arange(8)
        assert np.all(t["parent"][ch] == k) and np.all(t["depth"][ch] == t["depth"][k] + 1)
        assert np.array_equal(t["path"][ch], 2 * t["path"][k] + np.array([[a, b, c] for a in (0, 1) for b in (0, 1)
                                                                           for c in (0, 1)]))
    ls = t["level_start"]
    for d in range(len(ls) - 1):
        assert np.all(t["depth"][ls[d]:ls[d + 1]] == d)                      # BFS level order
    s = np.where(leaf & (t["src_num"] > 0))[0]
    assert np.all(np.diff(t["src_beg"][s]) == t["src_num"][s][:-1])           # leaf ranges in node order
    assert nn == ls[-1]


def _touch(t, a, b):
    da, db = t["depth"][a], t["depth"][b]
    md = max(da, db)
    ra, rb = 1 << (md - da), 1 << (md - db)
    ca = (2 * t["path"][a] + 1) * ra
    cb = (2 * t["path"][b] + 1) * rb
    dif = np.abs(ca - cb)
    return bool(np.all(dif <= ra + rb) and dif.max() == ra + rb)


def _colleagues(t, node):
    d = t["depth"][node]
    same = np.where(t["depth"] == d)[0]
    return [c for c in same if c != node and np.abs(t["path"][c] - t["path"][node]).max() <= 1]


@pytest.mark.parametrize("case", [(400, "ellipsoid", 4, 3), (300, "gaussmix", 3, 1), (600, "uniform", 5, 2)])
def test_lists_match_textbook_definitions(case):
    """U(B): leaves touching leaf B, and B itself.  V(B): children of the
    colleagues of B's parent not touching B.  W(B): descendants of B's
    colleagues whose parent touches B but which do not.  X(B): source
    leaves C coarser than B that touch B's parent but not B (the dual of W).
    All restricted to source-holding boxes."""
    n, dist, q, seed = case
    p = K.Problem.create(n, dist, q, np_=4, seed=seed)
    t = p.tree()
    nn = t["n_nodes"]
    leaf = t["child"] == -1
    src = (t["tag"] & 1) != 0
    trg = (t["tag"] & 2) != 0

    def desc(k):
        out, stack = [], [k]
        while stack:
            x = stack.pop()
            if t["child"][x] != -1:
                ch = list(range(t["child"][x], t["child"][x] + 8))
                out += ch
                stack += ch
        return out

    W = {b: set() for b in range(nn)}
    for b in range(nn):
        if not (leaf[b] and trg[b]):
            continue
        for c in _colleagues(t, b):
            for w in desc(c):
                if src[w] and not _touch(t, w, b) and _touch(t, t["parent"][w], b):
                    W[b].add(w)
    for b in range(nn):
        if not trg[b]:
            continue
        got = {L: set(t[f"{L}_idx"][t[f"{L}_ptr"][b]:t[f"{L}_ptr"][b + 1]].tolist()) for L in "uvwx"}
        U = set()
        if leaf[b]:
            U = {a for a in range(nn) if leaf[a] and src[a] and (a == b or _touch(t, a, b))}
        V = set()
        if t["parent"][b] >= 0:
            for c in _colleagues(t, t["parent"][b]) + [t["parent"][b]]:
                if t["child"][c] != -1:
                    V |= {v for v in range(t["child"][c], t["child"][c] + 8) if src[v] and not _touch(t, v, b) and v != b}
        X = set()
        if t["parent"][b] >= 0:
            X = {c for c in range(nn) if leaf[c] and src[c] and t["depth"][c] < t["depth"][b]
                 and not _touch(t, c, b) and _touch(t, c, t["parent"][b])}
        assert got["u"] == U, ("U", b)
        assert got["v"] == V, ("V", b)
        assert got["w"] == (W[b] if leaf[b] else set()), ("W", b)
        assert got["x"] == X, ("X", b)
        # X is the dual of W: C in X(B) implies B in W(C) when C is a target
        # leaf and B holds sources (W lists are SRCNODE-filtered)
        for c in X:
            if trg[c] and src[b]:
                assert b in W[c], ("X/W duality", b, c)


def test_generators():
    p = K.Problem.create(20000, "ellipsoid", 64, np_=4, seed=2).points()
    assert np.allclose(p["sx"] ** 2 + p["sy"] ** 2 + (p["sz"] / 0.25) ** 2, 1.0, atol=1e-12)
    u = K.Problem.create(20000, "uniform", 64, np_=4, seed=2).points()
    for k in ("sx", "sy", "sz", "tx"):
        assert u[k].min() >= -1.0 and u[k].max() < 1.0
    assert 0.0 <= u["den"].min() and u["den"].max() < 1.0
    g = K.Problem.create(20000, "gaussmix", 64, np_=4, seed=2).points()
    assert np.abs(g["sx"]).max() <= 1.0
    c = K.Problem.create(1000, "uniform", 64, np_=4, seed=2, coincide=1).points()
    assert np.array_equal(c["sx"], c["tx"])


def test_scatter_roundtrip_and_validate():
    p = K.Problem.create(5000, "uniform", 32, np_=4, seed=3, coincide=1)
    t = p.tree()
    pot_leaf = t["tx"] * 3.0 + t["ty"]
    pts = p.points()
    orig = p.scatter(pot_leaf)
    assert np.array_equal(orig, pts["tx"] * 3.0 + pts["ty"])
    # validate against a brute-force direct sum (r == 0 excluded)
    d = pts["tx"][:, None] - pts["sx"][None, :]
    r = np.sqrt(d ** 2 + (pts["ty"][:, None] - pts["sy"][None, :]) ** 2 + (pts["tz"][:, None] - pts["sz"][None, :]) ** 2)
    with np.errstate(divide="ignore"):
        inv = np.where(r > 0, 1.0 / r, 0.0)
    exact = (inv * pts["den"][None, :]).sum(1) / (4 * np.pi)
    err, ids, direct = p.validate(exact, 200)
    assert err < 1e-14 and np.allclose(direct, exact[ids], rtol=1e-14)


def test_from_points_matches_generated():
    a = K.Problem.create(3000, "uniform", 16, np_=4, seed=9)
    pts = a.points()
    b = K.Problem.from_points(pts["sx"], pts["sy"], pts["sz"], pts["den"], pts["tx"], pts["ty"], pts["tz"],
                              pts_max=16, np_=4)
    ta, tb = a.tree(), b.tree()
    for k in KEYS:
        assert np.array_equal(ta[k], tb[k]), k
    with pytest.raises(K.KifmmError):
        K.Problem.from_points(pts["sx"] * 3, pts["sy"], pts["sz"], pts["den"], pts_max=16, np_=4)


def test_point_file_roundtrip(tmp_path):
    """KIFMMPT1 files: separate targets and sources-are-targets; the tree
    from the file equals the tree from the same points in memory; a
    truncated file is an error, not a crash."""
    a = K.Problem.create(5000, "gaussmix", 16, np_=4, seed=12)
    pts = a.points()
    f = tmp_path / "p.bin"
    K.write_points(f, pts["sx"], pts["sy"], pts["sz"], pts["den"], pts["tx"], pts["ty"], pts["tz"])
    b = K.Problem.from_file(f, pts_max=16, np_=4)
    ta, tb = a.tree(), b.tree()
    for k in KEYS:
        assert np.array_equal(ta[k], tb[k]), k
    g = tmp_path / "c.bin"
    K.write_points(g, pts["sx"], pts["sy"], pts["sz"], pts["den"])
    c = K.Problem.from_file(g, pts_max=16, np_=4)
    assert c.view().coincident == 1 and np.array_equal(c.points()["tx"], pts["sx"])
    h = tmp_path / "t.bin"
    h.write_bytes(f.read_bytes()[:1000])
    with pytest.raises(K.KifmmError):
        K.Problem.from_file(h, pts_max=16, np_=4)


def test_too_many_coincident_points_is_an_error():
    """More than pts_max coincident sources can never be separated: the build
    stops at 30 levels with an error instead of splitting forever; exactly
    pts_max coincident sources stay in one leaf."""
    pts = np.zeros((3, 50)) + 0.3
    with pytest.raises(K.KifmmError, match="30 levels"):
        K.Problem.from_points(*pts, np.ones(50), pts_max=8, np_=4)
    t = K.Problem.from_points(*pts, np.ones(50), pts_max=50, np_=4).tree()
    assert t["n_nodes"] == 1 and t["src_num"][0] == 50


def test_operators_are_shared_per_np():
    """The operators depend on NP only and are built once per NP per
    process: problems of any size and distribution get identical arrays,
    and differ between NPs."""
    a = K.Problem.create(3000, "uniform", 32, np_=3, seed=1)
    b = K.Problem.create(5000, "ellipsoid", 16, np_=3, seed=2)
    c = K.Problem.create(3000, "uniform", 32, np_=4, seed=1)
    oa, ob, oc = a.operator_arrays(), b.operator_arrays(), c.operator_arrays()
    for k in ("uc2ue", "ue2uc", "dc2de", "de2dc"):
        assert np.array_equal(oa[k], ob[k]), k
        assert oa[k].shape != oc[k].shape, k
