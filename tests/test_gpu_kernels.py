"""Kernel-level GPU tests: each evaluator kernel against numpy / the oracle.

Pins the pieces the end-to-end parity test composes: the MFMA operand
layout, the FP64 GEMM, the reciprocal square root, the forward/inverse
pruned DFTs (FFTW r2c/c2r conventions: [z][y][x/2+1], unnormalised
inverse, fmm/trans.cc:271-416) and the 316 translation spectra
(fmm/omp_sse_block.cc:432-478).
"""
import numpy as np
import pytest

import kernel_lib as KL
import kifmm_amd as K
import oracle_lib as O

pytestmark = pytest.mark.gpu


def test_mfma_f64_operand_layout():
    lanes = np.arange(64)
    # A[i][k] in lane i + 16k, B[k][j] in lane j + 16k, D[i][j] in lane j + 16(i%4), element i//4
    rows = (lanes // 16)[:, None] + 4 * np.arange(4)[None, :]
    cols = np.repeat((lanes % 16)[:, None], 4, axis=1)
    out = KL.mfma_layout(lanes + 1.0, (lanes < 16).astype(float))   # D[i][j] = A[i][0] = i + 1
    assert np.array_equal(out, rows + 1.0), f"row map differs:\n{out}"
    out = KL.mfma_layout((lanes < 16).astype(float), lanes + 1.0)   # D[i][j] = B[0][j] = j + 1
    assert np.array_equal(out, cols + 1.0), f"column map differs:\n{out}"


def test_mfma_f32_operand_layout():
    """v_mfma_f32_16x16x4f32: A[i][k] in lane i + 16k, B[k][j] in lane j + 16k,
    D[i][j] in lane j + 16(i//4), element i%4 (the FP32 output layout differs
    from the FP64 one)."""
    lanes = np.arange(64)
    rows = 4 * (lanes // 16)[:, None] + np.arange(4)[None, :]
    cols = np.repeat((lanes % 16)[:, None], 4, axis=1)
    out = KL.mfma_layout_f32(lanes + 1.0, (lanes < 16).astype(float))   # D[i][j] = A[i][0] = i + 1
    assert np.array_equal(out, rows + 1.0), f"row map differs:\n{out}"
    out = KL.mfma_layout_f32((lanes < 16).astype(float), lanes + 1.0)   # D[i][j] = B[0][j] = j + 1
    assert np.array_equal(out, cols + 1.0), f"column map differs:\n{out}"


def test_mfma_bf16_operand_layout():
    """v_mfma_f32_16x16x32_bf16: A[i][k] in lane i + 16(k//8), element k%8;
    B[k][j] in lane j + 16(k//8), element k%8; D[i][j] in lane j + 16(i//4),
    element i%4.  Small integers are exact in bf16 and in the FP32 sums, so
    a random product pins the whole map."""
    rng = np.random.default_rng(17)
    A = rng.integers(-4, 5, (16, 32)).astype(np.float32)
    B = rng.integers(-4, 5, (32, 16)).astype(np.float32)
    lanes = np.arange(64)
    a = np.zeros((64, 8), np.float32)
    b = np.zeros((64, 8), np.float32)
    for l in lanes:
        for e in range(8):
            k = 8 * (l // 16) + e
            a[l, e] = A[l % 16, k]
            b[l, e] = B[k, l % 16]
    out = KL.mfma_layout_bf16(a, b)
    D = A @ B
    want = np.array([[D[4 * (l // 16) + r, l % 16] for r in range(4)] for l in lanes])
    assert np.array_equal(out, want), f"layout differs:\n{out}\n{want}"


@pytest.mark.parametrize("M,Kd,n", [(56, 152, 37), (152, 448, 20), (448, 56, 33), (26, 104, 5), (296, 1216, 17),
                                    (152, 160, 64), (152, 448, 9000), (56, 1216, 8200), (98, 56, 300),
                                    (1216, 152, 700), (208, 26, 40), (56, 152, 70001), (26, 56, 9000),
                                    (64, 160, 8192), (16, 8, 8200), (56, 150, 8193), (56, 152, 4096)])
def test_gemm_matches_numpy(M, Kd, n):
    """Small column counts take the split-K route (partials + ordered
    reduction), large ones the single-pass route, and short operators (<= 64
    rows, K <= 160 after padding to a multiple of 8) over >= 8192 columns the
    streaming kernel, column counts that are not multiples of 16 included."""
    rng = np.random.default_rng(M * 1000 + Kd)
    A = rng.standard_normal((M, Kd))
    B = rng.standard_normal((Kd, n))
    se = rng.integers(-6, 7, n)
    C0 = rng.standard_normal((M, n))
    ref = (A @ B) * np.ldexp(1.0, se)[None, :]
    got = KL.gemm(A, B, se)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    got = KL.gemm(A, B, se, C0, accumulate=True)
    assert np.max(np.abs(got - (C0 + ref))) <= 1e-13 * np.max(np.abs(C0 + ref))


def test_rsqrt_accuracy():
    rng = np.random.default_rng(3)
    x = np.concatenate([np.exp(rng.uniform(np.log(1e-30), np.log(1e30), 200000)),
                        rng.uniform(0.0, 12.0, 200000), [1.0, 4.0, 0.25, 2.0, 3.0]])
    y, ym = KL.rsqrt(x)
    exact = 1.0 / np.sqrt(x.astype(np.longdouble))
    relerr = np.abs((y.astype(np.longdouble) - exact) / exact).astype(np.float64)
    # v_rsq_f64 (< 2^-24) plus one Newton step: (3/2) e0^2 < 2^-47 (measured
    # raw estimate error 5.2e-8 gives ~4e-15)
    assert relerr.max() <= 2.0 ** -47, relerr.max()
    assert np.array_equal(y, ym)
    y0, ym0 = KL.rsqrt(np.array([0.0, 1.0]))
    assert ym0[0] == 0.0 and ym0[1] == 1.0


def _grid_from_surface(np_, vals):
    """scatter surface samples onto the (2np)^3 grid at offset np/2 (fmm/trans.cc:335-355,394-416)"""
    N = 2 * np_
    g = np.zeros((N, N, N))
    cnt = 0
    h = np_ // 2
    for i in range(np_):
        for j in range(np_):
            for k in range(np_):
                if i in (0, np_ - 1) or j in (0, np_ - 1) or k in (0, np_ - 1):
                    g[k + h, j + h, i + h] = vals[cnt]
                    cnt += 1
    return g


def _surface_from_grid(np_, g):
    out = []
    h = np_ // 2
    for i in range(np_):
        for j in range(np_):
            for k in range(np_):
                if i in (0, np_ - 1) or j in (0, np_ - 1) or k in (0, np_ - 1):
                    out.append(g[k + h, j + h, i + h])
    return np.array(out)


@pytest.mark.parametrize("np_", [3, 4, 6])
def test_forward_dft_matches_numpy_rfftn(np_):
    rng = np.random.default_rng(np_)
    nb, ne = 3, K.n_e(np_)
    den = rng.standard_normal((8 * nb, ne))
    depth = rng.integers(2, 6, 8 * nb)
    got = KL.dft_fwd(np_, den, depth)                   # [nb, F, 8]
    for b in range(nb):
        for o in range(8):
            g = _grid_from_surface(np_, den[8 * b + o] * 2.0 ** depth[8 * b + o])
            ref = np.fft.rfftn(g).reshape(-1)              # [kz][ky][kx] flattened
            err = np.max(np.abs(got[b, :, o] - ref)) / np.max(np.abs(ref))
            assert err <= 1e-14, (b, o, err)


@pytest.mark.parametrize("np_", [3, 4, 6])
def test_inverse_dft_matches_numpy_irfftn(np_):
    rng = np.random.default_rng(10 + np_)
    N, nb = 2 * np_, 2
    grids = rng.standard_normal((nb * 8, N, N, N))
    spec = np.stack([np.fft.rfftn(grids[k]).reshape(-1) for k in range(nb * 8)])   # [nb*8, F]
    spec = spec.reshape(nb, 8, -1).transpose(0, 2, 1).copy()                        # [nb, F, 8]
    got = KL.dft_inv(np_, spec)
    for k in range(nb * 8):
        ref = _surface_from_grid(np_, grids[k])
        assert np.max(np.abs(got[k] - ref)) <= 1e-13 * np.max(np.abs(ref)), k


@pytest.mark.parametrize("np_", [3, 4, 6])
def test_translation_spectra_match_oracle(np_):
    got = KL.translation_spectra(np_)                    # [F, 343]
    ref = O.translation_spectra(np_, O.KERNEL_EXACT)     # [343, N, N, H]
    ref = ref.reshape(343, -1).T
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    near = [(a + 3) + 7 * (b + 3) + 49 * (c + 3) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]
    assert np.all(got[:, near] == 0)


def test_rsq_clamp_f32():
    """min(1, 1/sqrt|x|) from v_rsq_f32 with the |x| and clamp modifiers, as
    the mixed U-list forms it: exactly 1 for |x| <= 1 (0, -0, tiny negative
    Gram rounding, inf), the raw estimate's accuracy (< 2^-22) beyond."""
    rng = np.random.default_rng(9)
    small = np.array([0.0, -0.0, 1e-30, -1e-30, -3e-5, 0.25, 0.999, 1.0, -1.0], dtype=np.float32)
    y = KL.rsq_clamp_f32(small)
    assert np.all(y == 1.0), y
    x = np.concatenate([np.exp(rng.uniform(0.0, np.log(1e30), 100000)), rng.uniform(1.0, 1e4, 100000),
                        -rng.uniform(1.0, 1e4, 1000)]).astype(np.float32)
    y = KL.rsq_clamp_f32(x).astype(np.float64)
    exact = np.minimum(1.0, 1.0 / np.sqrt(np.abs(x.astype(np.float64))))
    assert np.max(np.abs(y - exact) / exact) < 2.0 ** -22
    assert KL.rsq_clamp_f32(np.array([np.inf], dtype=np.float32))[0] == 0.0
