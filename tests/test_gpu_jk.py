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


@pytest.mark.parametrize("Q,n,na,nb,extra,seed", [(40, 90, 11, 7, 0, 8), (23, 37, 5, 0, 3, 9),
                                                  (31, 130, 9, 9, 0, 10), (17, 64, 0, 0, 2, 11),
                                                  (48, 600, 61, 55, 1, 12), (12, 50, 3, 6, 0, 13)])
def test_jk_rohf_matches_oracle(ctx, Q, n, na, nb, extra, seed):
    """Restricted open-shell J/Ka/Kb (occupied-factor form, one half
    transform for both spins) against the general-density oracle."""
    rng = np.random.default_rng(seed)
    L = syn.gen_L(Q, n, rng)
    m = max(na, nb)
    C = syn.gen_Cocc(n, max(m + extra, 1), rng)  # ldc = m + extra columns
    Pa, Pb = C[:, :na] @ C[:, :na].T, C[:, :nb] @ C[:, :nb].T
    J_ref, K_ref = oracle.jk(L, np.stack([Pa, Pb]))
    print(f"jk_rohf Q={Q} n={n} na={na} nb={nb} ldc={C.shape[1]}", flush=True)
    J, Ka, Kb = ctx.jk_rohf(*to_dev(L, C), na, nb)
    check_close(to_np(J), J_ref, 1e-12, "J_ref")
    check_close(to_np(Ka), K_ref[0], 1e-12, "Ka_ref")
    check_close(to_np(Kb), K_ref[1], 1e-12, "Kb_ref")
    if na == 0:
        assert not to_np(Ka).any() and not to_np(J).any()
    if nb == 0:
        assert not to_np(Kb).any()
