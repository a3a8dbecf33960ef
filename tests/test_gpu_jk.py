"""GPU parity of J/K and the AO->MO transforms against the FP64 oracle."""
import numpy as np
import pytest

import oracle
from conftest import check_close, to_dev, to_np
from mshqc_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _inputs(Q, n, o, seed):
    rng = np.random.default_rng(seed)
    L = syn.gen_L(Q, n, rng)
    C = syn.gen_Cocc(n, n, rng)  # full orthonormal basis: occupied | virtual
    return L, C[:, :o].copy(), C[:, o:].copy()


@pytest.mark.parametrize("Q,n,o,seed", [(23, 37, 5, 1), (48, 600, 60, 2), (5, 70, 9, 3)])
def test_transform_matches_oracle(ctx, Q, n, o, seed):
    L, Co, Cv = _inputs(Q, n, o, seed)
    B_ref = oracle.transform(L, Co, Cv)
    B = to_np(ctx.transform(*to_dev(L, Co, Cv)))
    print(f"transform Q={Q} n={n} o={o}", flush=True)
    check_close(B, B_ref, 1e-12, "B_ref")
    X_ref = oracle.half_transform(L, Co)
    X = to_np(ctx.half_transform(*to_dev(L, Co)))
    check_close(X, X_ref, 1e-12, "X_ref")


@pytest.mark.parametrize("Q,n,nspin,screen,seed", [(23, 37, 1, True, 4), (31, 50, 2, False, 5),
                                                   (40, 130, 1, False, 6)])
def test_jk_general_density_matches_oracle(ctx, Q, n, nspin, screen, seed):
    rng = np.random.default_rng(seed)
    L = syn.gen_L(Q, n, rng)
    if screen:
        L[3] = 0.0  # a vector the screening drops
        L[7] *= 1e-10
    P = rng.standard_normal((nspin, n, n))
    P = 0.5 * (P + P.transpose(0, 2, 1))
    P_arg = P[0] if nspin == 1 else P
    J_ref, K_ref = oracle.jk(L, P_arg, screen=screen)
    J, K = ctx.jk(*to_dev(L, P_arg), screen=screen)
    print(f"jk Q={Q} n={n} nspin={nspin} screen={screen}", flush=True)
    check_close(to_np(J), J_ref, 1e-12, "J_ref")
    check_close(to_np(K), K_ref, 1e-12, "K_ref")


def test_jk_factor_matches_general(ctx):
    rng = np.random.default_rng(7)
    Q, n, o = 40, 90, 11
    L = syn.gen_L(Q, n, rng)
    C = syn.gen_Cocc(n, o, rng)
    Y = np.sqrt(2.0) * C
    J_ref, K_ref = oracle.jk(L, Y @ Y.T)
    J, K = ctx.jk_factor(*to_dev(L, Y))
    check_close(to_np(J), J_ref, 1e-12, "J_ref")
    check_close(to_np(K), K_ref, 1e-12, "K_ref")
