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


# Split-BF16 transform bounds (max |B - B_ref| / max |B_ref|): each operand
# row block is cut at 2^-(9k-1) of its bound 2^e, which sits up to ~4x above
# a typical element, so a product of two operands carries ~16 x 2^-(9k-1)
# of its typical size; the transform is two such GEMMs.  k = 4 is then near
# 1e-9, where the FP32 rounding of the product classes past the first
# (2^-33 of a block sum) takes over.
TRANSFORM_TOL = {2: 1e-3, 3: 2e-6, 4: 1e-8}


@pytest.mark.parametrize("Q,n,o,seed", [(23, 37, 5, 11), (130, 300, 40, 12), (9, 257, 3, 13)])
@pytest.mark.parametrize("k", [2, 3, 4])
def test_transform_split_bf16(ctx, Q, n, o, seed, k):
    """mshqc_hip_transform with k split-BF16 slices against the oracle's FP64
    transform (cholesky_rmp2.cc:68-92), on ragged AO counts (37, 257 cross
    the 256-wide exponent blocks) and Q chunks."""
    L, Co, Cv = _inputs(Q, n, o, seed)
    B_ref = oracle.transform(L, Co, Cv)
    B = to_np(ctx.transform(*to_dev(L, Co, Cv), precision=k))
    err = np.abs(B - B_ref).max() / np.abs(B_ref).max()
    print(f"transform bf16x{k} Q={Q} n={n} o={o}: rel {err:.2e}", flush=True)
    assert err <= TRANSFORM_TOL[k]


def test_transform_split_bf16_energy(ctx):
    """The four-slice transform's effect on the CD-RMP2 energy it feeds.  The
    half transform is FP64; the virtual product cuts each row block of T and
    C_vir at 2^-35 of its bound, so B carries ~1e-9 relative error (max norm),
    random in sign.  The energy's error is a random sum over B's errors,
    measured at 1e-12 to 3e-12 relative at this shape (|E| ~ 4e3 Eh) and
    9e-14 to 1.5e-13 at o=100/v=1100/Q=4000: the test allows 1e-11, i.e.
    4e-8 Eh here, far below the 1 uEh bar."""
    L, Co, Cv = _inputs(400, 260, 20, 14)
    eo = np.sort(np.random.default_rng(15).uniform(-1.5, -0.3, 20))
    ev = 0.1 + 3.0 * (np.arange(240) / 240) ** 1.5
    dL, dCo, dCv, deo, dev = to_dev(L, Co, Cv, eo, ev)
    E64 = ctx.rmp2_energy(ctx.transform(dL, dCo, dCv), deo, dev)
    E4 = ctx.rmp2_energy(ctx.transform(dL, dCo, dCv, precision=4), deo, dev)
    print(f"energy from the four-slice transform: {E4!r} vs {E64!r} ({(E4 - E64) / E64:.1e} rel)", flush=True)
    assert abs(E4 - E64) <= 1e-11 * abs(E64)


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
