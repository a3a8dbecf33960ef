"""The oracle pinned against what the mathematics fixes: closed forms,
invariants, finite differences and the repository's own OMP2 density
formulas (no golden vectors exist in the reference's tests)."""
import numpy as np
import pytest

import oracle
from oracle import omp2_ref
from mshqc_amd import synthetic as syn


def test_q1_o1_closed_form():
    """Q = 1, o = 1: G_ab = b_a b_b, E = sum_ab b_a^2 b_b^2 / D."""
    rng = np.random.default_rng(1)
    v = 7
    b = rng.standard_normal(v)
    eo = np.array([-0.7])
    ev = np.linspace(0.2, 1.3, v)
    E = oracle.rmp2_energy(b[None, None, :], eo, ev)
    D = 2 * eo[0] - ev[:, None] - ev[None, :]
    assert abs(E - np.sum(np.outer(b * b, b * b) / D)) < 1e-14


def test_rotation_invariance_over_q():
    B, eo, ev = syn.rmp2_inputs(12, 3, 8, 2)
    R, _ = np.linalg.qr(np.random.default_rng(3).standard_normal((12, 12)))
    Br = np.einsum("pq,qia->pia", R, B)
    E1, g1, do1, dv1 = oracle.rmp2_grad(B, eo, ev)
    E2, g2, do2, dv2 = oracle.rmp2_grad(Br, eo, ev)
    assert abs(E1 - E2) < 1e-13 * abs(E1)
    assert np.allclose(np.einsum("pq,qia->pia", R, g1), g2, rtol=0, atol=1e-13 * np.abs(g1).max())
    assert np.allclose(do1, do2, atol=1e-14) and np.allclose(dv1, dv2, atol=1e-14)


def test_einsum_restatement_of_energy():
    B, eo, ev = syn.rmp2_inputs(9, 3, 6, 4)
    g = np.einsum("qia,qjb->ijab", B, B)
    D = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    E = np.sum(g * (2 * g - g.transpose(0, 1, 3, 2)) / D)
    assert abs(oracle.rmp2_energy(B, eo, ev) - E) < 1e-14


@pytest.mark.parametrize("seed", [5, 6])
def test_gradient_identities_and_fd(seed):
    B, eo, ev = syn.rmp2_inputs(20, 4, 9, seed)
    E, g, deo, dev = oracle.rmp2_grad(B, eo, ev)
    assert abs(np.sum(B * g) - 4 * E) < 1e-13 * abs(E)        # degree-4 homogeneity in B
    assert abs(deo.sum() + dev.sum()) < 1e-14                   # E depends on eps via D only
    assert abs(eo @ deo + ev @ dev + E) < 1e-13 * abs(E)        # degree -1 in eps
    h = 1e-5
    rng = np.random.default_rng(seed)
    for _ in range(6):
        idx = tuple(rng.integers(0, s) for s in B.shape)
        Bp, Bm = B.copy(), B.copy()
        Bp[idx] += h
        Bm[idx] -= h
        fd = (oracle.rmp2_energy(Bp, eo, ev) - oracle.rmp2_energy(Bm, eo, ev)) / (2 * h)
        assert abs(fd - g[idx]) < 1e-8 * max(1.0, abs(g[idx]))


def test_t2_and_energy_from_t2():
    B, eo, ev = syn.rmp2_inputs(10, 3, 5, 7)
    E, t2 = oracle.rmp2_energy(B, eo, ev, want_t2=True)
    D = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    # E = sum t (2 g - g^T) with g = t D
    g = t2 * D
    assert abs(E - np.sum(t2 * (2 * g - g.transpose(0, 1, 3, 2)))) < 1e-14


def test_ump2_equals_rmp2_for_identical_spins():
    B, eo, ev = syn.rmp2_inputs(15, 3, 7, 8)
    Eu, grads = oracle.ump2(B, B, eo, ev, eo, ev, grad=True)
    E, g, deo, dev = oracle.rmp2_grad(B, eo, ev)
    assert abs(sum(Eu) - E) < 1e-14 * abs(E) * 10
    assert np.allclose(grads[0] + grads[1], g, atol=1e-15 * 100)
    assert np.allclose(grads[2] + grads[4], deo, atol=1e-15 * 100)
    assert np.allclose(grads[3] + grads[5], dev, atol=1e-15 * 100)


def test_omp2_energy_and_opdm_diagonals_pin_ump2_backward():
    """dE/deps from the backward equals the diagonal of the repository's own
    OMP2 one-particle density (cholesky_omp2.cc:232-326), and the OMP2
    energy-from-t2 equals the direct UMP2 sum (cholesky_omp2.cc:189-230)."""
    (Ba, eoa, eva), (Bb, eob, evb) = syn.ump2_inputs(14, 3, 6, 2, 7, 9)
    E, (dBa, dBb, deoa, deva, deob, devb) = oracle.ump2(Ba, Bb, eoa, eva, eob, evb, grad=True)
    taa, tbb, tab = omp2_ref.t2_blocks(Ba, Bb, eoa, eva, eob, evb)
    assert abs(omp2_ref.energy(taa, tbb, tab, eoa, eva, eob, evb) - sum(E)) < 1e-14 * 10
    Goo_a, Gvv_a, Goo_b, Gvv_b = omp2_ref.opdm(taa, tbb, tab)
    assert np.allclose(np.diag(Goo_a), deoa, atol=1e-15 * 10)
    assert np.allclose(np.diag(Gvv_a), deva, atol=1e-15 * 10)
    assert np.allclose(np.diag(Goo_b), deob, atol=1e-15 * 10)
    assert np.allclose(np.diag(Gvv_b), devb, atol=1e-15 * 10)
    # RMP2: dE/deps_p = G_alpha(p,p) + G_beta(p,p) with identical spins
    B, eo, ev = syn.rmp2_inputs(14, 3, 6, 10)
    _, _, deo, dev = oracle.rmp2_grad(B, eo, ev)
    taa, tbb, tab = omp2_ref.t2_blocks(B, B, eo, ev, eo, ev)
    Goo_a, Gvv_a, Goo_b, Gvv_b = omp2_ref.opdm(taa, tbb, tab)
    assert np.allclose(np.diag(Goo_a) + np.diag(Goo_b), deo, atol=1e-14)
    assert np.allclose(np.diag(Gvv_a) + np.diag(Gvv_b), dev, atol=1e-14)


def test_jk_brute_force():
    rng = np.random.default_rng(11)
    Q, n = 6, 5
    L = syn.gen_L(Q, n, rng)
    eri = np.einsum("qij,qkl->ijkl", L, L)
    P = rng.standard_normal((n, n))
    P = P + P.T
    J, K = oracle.jk(L, P)
    assert np.allclose(J, np.einsum("ijkl,kl->ij", eri, P), atol=1e-14)
    assert np.allclose(K, np.einsum("ikjl,kl->ij", eri, P), atol=1e-14)
    L[2] *= 1e-10  # screened out of K only
    J2, K2 = oracle.jk(L, P, screen=True)
    Lk = L.copy()
    Lk[2] = 0
    assert np.allclose(K2, np.einsum("ikjl,kl->ij", np.einsum("qij,qkl->ijkl", Lk, Lk), P), atol=1e-14)


def test_transform_and_half_transform():
    rng = np.random.default_rng(12)
    L = syn.gen_L(4, 6, rng)
    C = rng.standard_normal((6, 6))
    B = oracle.transform(L, C[:, :2], C[:, 2:])
    assert np.allclose(B, np.einsum("mi,qmn,na->qia", C[:, :2], L, C[:, 2:]), atol=1e-14)
    X = oracle.half_transform(L, C[:, :2])
    assert np.allclose(X, np.einsum("qmn,ni->qmi", L, C[:, :2]), atol=1e-14)


def test_cholesky_exact_and_threshold():
    rng = np.random.default_rng(13)
    n = 4
    A = rng.standard_normal((n * n, 7))
    M = A @ A.T
    M = 0.5 * (M + M.T)
    eri = M.reshape(n, n, n, n)
    L = oracle.cholesky(eri, 1e-12)
    assert L.shape[0] == 7
    assert np.allclose(L.T @ L, M, atol=1e-10)
    L2 = oracle.cholesky(eri, 1e-2)
    d = np.diag(M - L2.T @ L2)
    assert d.max() < 1e-2
    assert oracle.cholesky(eri, 1e9).shape[0] == 0


def test_rank1_reference_loop_equals_energy():
    """The timing baseline's restatement of the reference loop (rank-1 updates,
    cholesky_rmp2.cc:118-155) over all ordered pairs equals the oracle energy,
    and over a subset equals the sum of those pairs' energies."""
    B, eo, ev = syn.rmp2_inputs(40, 5, 23, 17)
    E = oracle.rmp2_energy(B, eo, ev)
    o = B.shape[1]
    allp = np.array([(i, j) for i in range(o) for j in range(o)])
    assert abs(oracle.rmp2_energy_pairs_rank1(B, eo, ev, allp) - E) <= 1e-12 * abs(E)
    sub = allp[[1, 7, 12]]
    want = sum(oracle.rmp2_grad_pairs(B, eo, ev, sub[k:k + 1])[0] for k in range(len(sub)))
    assert abs(oracle.rmp2_energy_pairs_rank1(B, eo, ev, sub) - want) <= 1e-12 * abs(want)
