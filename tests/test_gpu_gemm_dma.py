"""GPU parity of the FP64 GEMM engine's direct-to-LDS kernels (gemm_f64.hip)
through the entry points that use them, against the FP64 oracle.

The shapes are chosen to reach each kernel and each edge of its work split:
  - 128 x 128 tiles with 8-deep stages (general-density K, n = 518: the
    contraction 518 ends in a partial stage; the batch-reduced lower-half
    product runs its flattened (vector, k) range across stage boundaries);
  - 128 x 128 tiles with 16-deep stages and a split contraction (the FP64
    RMP2 gradient at Q = 520, one occupied block: a single GEMM with two
    contraction layers summed afterwards);
  - 256 x 64 tiles: the half transforms L^q C (at most 64 columns) of all
    vectors as one tall GEMM, with a partial last row tile, storing each
    X_q transposed (the FP64 transform) or not (the factored exchange);
  - 128 x 128 tiles for tall outputs of 65-128 columns (o = 96);
  - 64 x 64 tiles: shorter half transforms (n x o), batch-reduced products
    with short per-vector contractions (K = 62 factor columns, 300 x 300
    output), and a split single GEMM (Q = 130);
  - the register-staged kernel: odd dimensions, where 16-byte segments
    would leave a row (covered by the ragged shapes of test_gpu_jk.py).
References: cholesky_rhf.cc:99-137 (J/K), cholesky_rmp2.cc:68-92 (transform),
the oracle's RMP2 gradient for the backward contractions.
"""
import numpy as np
import pytest

import oracle
from conftest import check_close, to_dev, to_np
from mshqc_amd import synthetic as syn

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("Q,n,nspin,seed", [(19, 518, 2, 21), (60, 300, 1, 22)])
def test_jk_general_density_dma_tiles(ctx, Q, n, nspin, seed):
    """n = 518: M = N >= 512 puts both products on 128 x 128 tiles with a
    partial last stage (518 = 64 * 8 + 6); n = 300: the batched product on
    128 x 128 tiles (its tiles over 60 vectors fill the chip), the reduced
    one on 64 x 64 tiles split over contraction layers."""
    rng = np.random.default_rng(seed)
    L = syn.gen_L(Q, n, rng)
    P = rng.standard_normal((nspin, n, n))
    P = 0.5 * (P + P.transpose(0, 2, 1))
    P_arg = P[0] if nspin == 1 else P
    print(f"jk dma Q={Q} n={n} nspin={nspin}", flush=True)
    J_ref, K_ref = oracle.jk(L, P_arg)
    J, K = ctx.jk(*to_dev(L, P_arg))
    check_close(to_np(J), J_ref, 1e-12, "J_ref")
    check_close(to_np(K), K_ref, 1e-12, "K_ref")


def test_jk_factor_short_contractions(ctx):
    """K = sum_q (L^q Y)(L^q Y)^T with 62 factor columns: the reduced product
    contracts 62 values per vector, so stages of 16 straddle vectors."""
    rng = np.random.default_rng(23)
    Q, n, m = 45, 300, 62
    L = syn.gen_L(Q, n, rng)
    Y = syn.gen_Cocc(n, m, rng)
    print(f"jk_factor dma Q={Q} n={n} m={m}", flush=True)
    J_ref, K_ref = oracle.jk(L, Y @ Y.T)
    J, K = ctx.jk_factor(*to_dev(L, Y))
    check_close(to_np(J), J_ref, 1e-12, "J_ref")
    check_close(to_np(K), K_ref, 1e-12, "K_ref")


def test_transform_dma_tiles(ctx):
    """B = C_occ^T L C_vir at n = 518, o = 30: the half transform of all 21
    vectors as one GEMM of 10878 rows on 64 x 64 tiles with a partial last
    stage, stored transposed, then the virtual product as one GEMM."""
    rng = np.random.default_rng(24)
    Q, n, o = 21, 518, 30
    L = syn.gen_L(Q, n, rng)
    C = syn.gen_Cocc(n, n, rng)
    Co, Cv = C[:, :o].copy(), C[:, o:].copy()
    print(f"transform dma Q={Q} n={n} o={o}", flush=True)
    B_ref = oracle.transform(L, Co, Cv)
    B = to_np(ctx.transform(*to_dev(L, Co, Cv)))
    check_close(B, B_ref, 1e-12, "B_ref")


@pytest.mark.parametrize("Q,n,o,seed", [(60, 302, 58, 31), (40, 256, 96, 32)])
def test_transform_tall_tiles(ctx, Q, n, o, seed):
    """The half transform as one tall GEMM storing each X_q transposed:
    Q*n = 18120 rows of 58 columns on 256 x 64 tiles (70 whole tiles and a
    partial one of 200 rows), and 10240 rows of 96 columns on 128 x 128 tiles;
    the second product then runs as one (Q*o) x v x n GEMM."""
    rng = np.random.default_rng(seed)
    L = syn.gen_L(Q, n, rng)
    C = syn.gen_Cocc(n, n, rng)
    Co, Cv = C[:, :o].copy(), C[:, o:].copy()
    print(f"transform tall tiles Q={Q} n={n} o={o}", flush=True)
    B_ref = oracle.transform(L, Co, Cv)
    B = to_np(ctx.transform(*to_dev(L, Co, Cv)))
    check_close(B, B_ref, 1e-12, "B_ref")
    X_ref = oracle.half_transform(L, Co)
    X = to_np(ctx.half_transform(*to_dev(L, Co)))
    check_close(X, X_ref, 1e-12, "X_ref")


def test_jk_factor_tall_tiles(ctx):
    """The factored exchange's half transforms L^q Y over 60 vectors of
    n = 300 as one GEMM of 18000 rows on 256 x 64 tiles (50 columns)."""
    rng = np.random.default_rng(33)
    Q, n, m = 60, 300, 50
    L = syn.gen_L(Q, n, rng)
    Y = syn.gen_Cocc(n, m, rng)
    print(f"jk_factor tall tiles Q={Q} n={n} m={m}", flush=True)
    J_ref, K_ref = oracle.jk(L, Y @ Y.T)
    J, K = ctx.jk_factor(*to_dev(L, Y))
    check_close(to_np(J), J_ref, 1e-12, "J_ref")
    check_close(to_np(K), K_ref, 1e-12, "K_ref")


@pytest.mark.parametrize("Q,o,v,seed", [(520, 4, 300, 25), (130, 4, 260, 26)])
def test_rmp2_grad_split_contraction(ctx, Q, o, v, seed):
    """FP64 RMP2 gradient whose dE/dB contraction is one GEMM with few output
    tiles and a long contraction (o v = 1200 / 1040), which the engine splits
    into contraction layers: Q = 520 on 128 x 128 tiles, Q = 130 on 64 x 64."""
    B, eo, ev = syn.rmp2_inputs(Q, o, v, seed)
    print(f"rmp2_grad dma Q={Q} o={o} v={v}", flush=True)
    E_ref, g_ref, deo_ref, dev_ref = oracle.rmp2_grad(B, eo, ev)
    E, g, d_o, d_v = ctx.rmp2_grad(*to_dev(B, eo, ev))
    assert abs(E - E_ref) <= 1e-12 * abs(E_ref)
    check_close(to_np(g), g_ref, 1e-12, "dE/dB")
    check_close(to_np(d_o), deo_ref, 1e-12, "dE/deps_o")
    check_close(to_np(d_v), dev_ref, 1e-12, "dE/deps_v")
