"""The C ABI: every declared entry point is exported, errors are returned
(never exit), and the host-side bounds audit rejects malformed inputs."""
import ctypes as C
import os
import re
from pathlib import Path

import numpy as np
import pytest

import kifmm_amd as K
import oracle_lib as O

ROOT = Path(__file__).resolve().parents[1]


def declared(header):
    text = Path(header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:kifmm|oracle)_\w+)\s*\(", text)))


@pytest.mark.parametrize("header", ["include/kifmm_host.h", "include/kifmm_hip.h", "include/kifmm_mlp.h"])
def test_library_exports_every_declared_symbol(header):
    names = declared(ROOT / header)
    assert names
    missing = [n for n in names if not hasattr(K.lib, n)]
    assert not missing, missing


def test_kernel_probes_live_in_the_test_library_only():
    """include/kifmm_kernels.h is exported by build/libkifmm_ktest.so, and
    none of its probes ships in the product library."""
    import kernel_lib as KL

    names = declared(ROOT / "include/kifmm_kernels.h")
    assert names
    assert not [n for n in names if not hasattr(KL.lib, n)]
    assert not [n for n in names if hasattr(K.lib, n)]


def test_oracle_exports_every_declared_symbol():
    names = declared(ROOT / "oracle" / "kifmm_oracle.h")
    assert not [n for n in names if not hasattr(O.lib, n)]


def test_invalid_inputs_are_errors():
    with pytest.raises(K.KifmmError) as e:
        K.Problem.create(1000, "nonsense", 64, np_=4)
    assert "invalid distribution" in str(e.value)
    with pytest.raises(K.KifmmError):
        K.Problem.create(0, "uniform", 64, np_=4)
    p = K.Problem.create(4096, "uniform", 64, np_=5, seed=1)
    with pytest.raises(K.KifmmError) as e:
        K.gpu_plan(p)
    assert e.value.code == -4


def _view_with(p, field, arr):
    v = p.view()
    setattr(v, field, arr.ctypes.data_as(C.POINTER(C.c_int)))
    return v


def test_plan_rejects_corrupted_views():
    """Fault injection from the host: indices a kernel would follow out of
    bounds are caught by kifmm_gpu_plan (which kifmm_gpu_create runs first)."""
    p = K.Problem.create(20000, "ellipsoid", 32, np_=4, seed=3)
    t = p.tree()
    assert K.gpu_plan(p)["u_pairs"] > 0
    for field, mutate in (
        ("child", lambda a: a.__setitem__(np.where(a != -1)[0][0], len(a) + 5)),
        ("u_idx", lambda a: a.__setitem__(3, 10 ** 7)),
        ("trg_num", lambda a: a.__setitem__(np.where(a > 0)[0][-1], 10 ** 8)),
        ("path", lambda a: a.__setitem__(5 * 3, -1)),
        ("depth", lambda a: a.__setitem__(9, 40)),
    ):
        arr = np.ascontiguousarray(t[field].reshape(-1).copy(), dtype=np.int32)
        mutate(arr)
        v = _view_with(p, field, arr)
        w = K.WorkCounts()
        rc = K.lib.kifmm_gpu_plan(C.byref(v), C.byref(p.operators()), C.byref(w))
        assert rc == -1, field
        assert "malformed" in K.last_error() or "internal" in K.last_error()


def test_work_counts_match_reference_estimators():
    """u_pairs and the V count follow the reference's flop estimators
    (fmm/driver.cc:44-85): 11 flops per U pair, 4*eff per V entry."""
    p = K.Problem.create(65536, "uniform", 64, np_=4, seed=0, coincide=1)
    t = p.tree()
    w = K.gpu_plan(p)
    pairs = sum(int(t["trg_num"][i]) * int(t["src_num"][s]) for i in range(t["n_nodes"])
                for s in t["u_idx"][t["u_ptr"][i]:t["u_ptr"][i + 1]])
    assert w["u_pairs"] == pairs
    assert w["v_entries"] == t["v_ptr"][-1] == 640584


def test_plan_pairs_sibling_leaves():
    """The mixed U-list takes two sibling leaves per work item when their
    targets fill at most 10 groups of 16 (each leaf its own groups): every
    leaf of a uniform 16-point-per-leaf tree is in a pair, none of a
    128-point-per-leaf one."""
    for q, want in ((64, 0.5), (256, 0.0)):
        p = K.Problem.create(65536, "uniform", q, np_=4, seed=0)
        t = p.tree()
        leaves = int(np.sum((t["child"] == -1) & (t["trg_num"] > 0)))
        assert K.gpu_plan(p)["pair_items"] == want * leaves, q


@pytest.mark.skipif(K.gpu_available(), reason="checks the no-device error path")
def test_create_without_device_returns_error():
    p = K.Problem.create(4096, "uniform", 64, np_=4, seed=1)
    with pytest.raises(K.KifmmError) as e:
        K.GpuEvaluator(p)
    assert e.value.code == -3


def test_environment_surface(monkeypatch):
    for k in ("NP", "NV", "C", "SEED", "BS", "UPBS"):
        monkeypatch.delenv(k, raising=False)
    assert (K.lib.kifmm_getenv_accuracy(), K.lib.kifmm_getenv_validate(), K.lib.kifmm_getenv_coincide(),
            K.lib.kifmm_getenv_seed(), K.lib.kifmm_getenv_block_size(), K.lib.kifmm_getenv_upbs()) == (6, 1000, 0, 26, 49, 1000)
    monkeypatch.setenv("NP", "4")
    monkeypatch.setenv("C", "yes")
    monkeypatch.setenv("SEED", "3")
    monkeypatch.setenv("NV", "-2")                 # invalid -> default, as fmm/util.c:176-193
    assert K.lib.kifmm_getenv_accuracy() == 4 and K.lib.kifmm_getenv_coincide() == 1
    assert K.lib.kifmm_getenv_seed() == 3 and K.lib.kifmm_getenv_validate() == 1000
    p = K.Problem.create(3000, "uniform", 32)      # NP, C, SEED from the environment
    assert p.view().np == 4 and p.view().coincident == 1
    q = K.Problem.create(3000, "uniform", 32, np_=4, seed=3, coincide=1)
    assert np.array_equal(p.points()["sx"], q.points()["sx"])
