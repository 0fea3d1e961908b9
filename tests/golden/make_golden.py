"""Writes the small fixtures under tests/golden/ (run from the repository root).

The reference holds no golden vectors, so the fixtures come from sources
independent of both the oracle and the product: glibc's drand48, numpy's
SVD and FFT.  oracle_small.npz is a regression fixture of the oracle itself
(pinned by the independent checks in test_oracle.py).
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
import oracle_lib as O  # noqa: E402

OOFP = 1.0 / (4.0 * np.pi)


def kmat(src, trg):
    d = src[:, :, None] - trg[:, None, :]
    return OOFP / np.sqrt((d ** 2).sum(0))


def main():
    libc = C.CDLL("libc.so.6")
    libc.drand48.restype = C.c_double
    libc.srand48.argtypes = [C.c_long]
    dr = {}
    for seed in (0, 1, 2, 3, 26):
        libc.srand48(seed)
        dr[f"seed{seed}"] = np.array([libc.drand48() for _ in range(64)])
    np.savez(HERE / "drand48_glibc.npz", **dr)

    np_ = 3
    eps = 0.1 ** (np_ + 1)
    ue, uc, de, dc = O.surface(np_, 1.0), O.surface(np_ + 2, 3.0), O.surface(np_, 3.0), O.surface(np_, 1.0)
    out = {}
    for name, (s, t) in {"uc2ue": (ue, uc), "dc2de": (de, dc)}.items():
        Km = kmat(s, t).T
        U, sv, Vt = np.linalg.svd(Km, full_matrices=False)
        r = int((sv >= eps * sv[0]).sum())
        out[name] = (Vt[:r].T / sv[:r]) @ U[:, :r].T
        out[name + "_sigma"] = sv
    np.savez(HERE / "pinv_np3_numpy.npz", **out)

    rng = np.random.default_rng(2024)
    g = rng.standard_normal((6, 6, 6))
    np.savez(HERE / "rfftn_N6_numpy.npz", grid=g, spectrum=np.fft.rfftn(g))

    o = O.Oracle(2000, "uniform", 16, 4, 0, 1)
    o.build_operators(O.KERNEL_EXACT)
    pot, _ = o.evaluate(O.KERNEL_EXACT)
    err, ids, direct = o.validate(pot, 100)
    np.savez(HERE / "oracle_small.npz", pot=pot, err=np.array(err), ids=ids, direct=direct)


if __name__ == "__main__":
    main()
