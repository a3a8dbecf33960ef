"""CD-OMP2: the oracle's restatement of the reference loop
(src/mp2/cholesky_omp2.cc:331-513) pinned by what the mathematics fixes, and
the host matrix exponential used by the orbital rotation.  Runs on the CPU
with the build's own integral engine and host Cholesky decomposition."""
import os

import numpy as np
import pytest
import scipy.linalg as sl

import oracle
from conftest import ROOT
from oracle import omp2_ref, scf_ref

BD = os.path.join(ROOT, "mshqc_amd", "data", "basis")


@pytest.fixture(scope="module")
def m():
    from mshqc_amd import native_module

    return native_module()


def system(m, atoms, basis, thr=1e-8, mult=1):
    mol = m.Molecule(0, mult)
    for Z, x, y, z in atoms:
        mol.add_atom(Z, x, y, z)
    b = m.BasisSet(basis, mol, BD)
    eng = m.IntegralEngine(mol, b)
    chol = m.CholeskyERI(thr)
    chol.decompose(eng.compute_eri())
    n = b.n_basis_functions()
    L = np.array(chol.get_L_vectors()).reshape(-1, n, n)
    S = eng.compute_overlap()
    H = eng.compute_kinetic() + eng.compute_nuclear()
    return mol, S, H, L


def rhf_guess(mol, S, H, L, nocc):
    r = scf_ref.cd_rhf(S, H, L, nocc, mol.nuclear_repulsion_energy(), energy_threshold=1e-11, max_iterations=300)
    return dict(C_a=r["C"], C_b=r["C"], P_a=0.5 * r["P"], P_b=0.5 * r["P"], eps_a=r["eps"], eps_b=r["eps"],
                e_scf=r["energy"])


def uhf_guess(mol, S, H, L, na, nb):
    r = scf_ref.cd_uhf(S, H, L, na, nb, mol.nuclear_repulsion_energy(), energy_threshold=1e-11,
                       max_iterations=300)
    return dict(C_a=r["Ca"], C_b=r["Cb"], P_a=r["Pa"], P_b=r["Pb"], eps_a=r["ea"], eps_b=r["eb"],
                e_scf=r["energy"])


def test_first_iteration_is_rmp2(m):
    """RHF orbitals: the first iteration's amplitude energy is the CD-RMP2
    energy, whose oracle matches the reference's compiled CholeskyRMP2."""
    mol, S, H, L = system(m, [(1, 0.0, 0.0, 1.4 * k) for k in range(4)], "cc-pvdz")
    g = rhf_guess(mol, S, H, L, 2)
    r = omp2_ref.cd_omp2(H, L, g["C_a"], g["C_b"], g["P_a"], g["P_b"], g["eps_a"], g["eps_b"], 2, 2, g["e_scf"],
                         max_iterations=3)
    C, eps = g["C_a"], g["eps_a"]
    e_rmp2 = oracle.rmp2_energy(oracle.transform(L, C[:, :2], C[:, 2:]), eps[:2], eps[2:])
    assert abs(r["trace"][0][1] - e_rmp2) < 1e-12 * abs(e_rmp2)
    assert abs(r["trace"][0][0] - (g["e_scf"] + e_rmp2)) < 1e-12
    # the loop lowers the total below the MP2 one
    assert r["trace"][-1][0] < r["trace"][0][0]
    # converged: the total on exit is the one before the last iteration
    r = omp2_ref.cd_omp2(H, L, g["C_a"], g["C_b"], g["P_a"], g["P_b"], g["eps_a"], g["eps_b"], 2, 2, g["e_scf"])
    assert r["converged"] and r["energy_total"] == r["trace"][-2][0]


def test_first_iteration_is_ump2_and_orbitals_stay_orthonormal(m):
    """UHF orbitals (H3 chain doublet, 2 alpha / 1 beta, cc-pVDZ): iteration 1
    gives the CD-UMP2 components; rotated orbitals stay S-orthonormal."""
    mol, S, H, L = system(m, [(1, 0.0, 0.0, 1.4 * k) for k in range(3)], "cc-pvdz", mult=2)
    g = uhf_guess(mol, S, H, L, 2, 1)
    r = omp2_ref.cd_omp2(H, L, g["C_a"], g["C_b"], g["P_a"], g["P_b"], g["eps_a"], g["eps_b"], 2, 1, g["e_scf"],
                         max_iterations=6, energy_threshold=1e-12, gradient_threshold=1e-10)
    Ca, Cb, ea, eb = g["C_a"], g["C_b"], g["eps_a"], g["eps_b"]
    E = oracle.ump2(oracle.transform(L, Ca[:, :2], Ca[:, 2:]), oracle.transform(L, Cb[:, :1], Cb[:, 1:]),
                    ea[:2], ea[2:], eb[:1], eb[1:])
    assert abs(r["trace"][0][1] - sum(E)) < 1e-12 * abs(sum(E))
    assert r["iterations"] == 6 and not r["converged"]
    for C in (r["C_a"], r["C_b"]):
        assert np.abs(C.T @ S @ C - np.eye(C.shape[1])).max() < 1e-12
    assert not np.allclose(r["C_a"], Ca)  # the orbitals did move
    assert r["energy_total"] == r["trace"][-1][0]  # not converged: the last iteration's total


def test_one_electron_stops_at_scf(m):
    """H atom (1 alpha, 0 beta): no pairs, so G = 0 and w = 2 F_vo is the
    converged SCF orbital gradient (zero): the loop stops at iteration 1 with
    the SCF energy."""
    mol, S, H, L = system(m, [(1, 0.0, 0.0, 0.0)], "cc-pvdz", mult=2)
    g = uhf_guess(mol, S, H, L, 1, 0)
    r = omp2_ref.cd_omp2(H, L, g["C_a"], g["C_b"], g["P_a"], g["P_b"], g["eps_a"], g["eps_b"], 1, 0, g["e_scf"])
    assert r["converged"] and r["iterations"] == 1
    assert r["e_mp2"] == 0.0 and r["energy_total"] == g["e_scf"]
    assert r["trace"][0][2] < 1e-9


def _rotated(C, o, c, k, h):
    # C expm(K): occupied k gains h C_c, virtual c loses h C_k (to first order)
    n = C.shape[0]
    K = np.zeros((n, n))
    K[o + c, k] = h
    K[k, o + c] = -h
    return C @ sl.expm(K)


@pytest.mark.parametrize("Q,n,na,nb,seed", [(30, 10, 3, 2, 3), (12, 7, 1, 1, 4), (25, 9, 4, 0, 5)])
def test_two_particle_grad_finite_differences(Q, n, na, nb, seed):
    """The two-particle term is the MP2 energy's derivative under the orbital
    rotation at fixed denominators: central differences of the energy of
    B(C expm(kappa)) (step 1e-4, error ~h^2) against the analytic term, for
    entries of both spins; a dropped term or a transposed block fails."""
    from mshqc_amd import synthetic as syn

    rng = np.random.default_rng(seed)
    L = syn.gen_L(Q, n, rng)
    Ca = sl.qr(rng.standard_normal((n, n)))[0]
    Cb = sl.qr(rng.standard_normal((n, n)))[0]
    eoa, eva = syn.gen_eps(na, n - na, rng)
    eob, evb = syn.gen_eps(nb, n - nb, rng)

    def B(C, o):
        return oracle.transform(L, C[:, :o], C[:, o:]) if o else np.zeros((Q, 0, n))

    def E(Ca, Cb):
        return sum(oracle.ump2(B(Ca, na), B(Cb, nb), eoa, eva, eob, evb))

    _, (dBa, dBb, *_) = oracle.ump2(B(Ca, na), B(Cb, nb), eoa, eva, eob, evb, grad=True)
    h = 1e-4
    for spin, C, o, dB in ((0, Ca, na, dBa), (1, Cb, nb, dBb)):
        if o == 0:
            continue
        w2 = omp2_ref.two_particle_grad(dB, oracle.transform(L, C[:, :o], C[:, :o]),
                                        oracle.transform(L, C[:, o:], C[:, o:]))
        assert w2.shape == (n - o, o)
        for c, k in {(0, 0), (n - o - 1, o - 1), (min(2, n - o - 1), min(1, o - 1))}:
            if spin == 0:
                fd = (E(_rotated(Ca, o, c, k, h), Cb) - E(_rotated(Ca, o, c, k, -h), Cb)) / (2 * h)
            else:
                fd = (E(Ca, _rotated(Cb, o, c, k, h)) - E(Ca, _rotated(Cb, o, c, k, -h))) / (2 * h)
            assert abs(fd - w2[c, k]) <= 1e-6 * abs(w2[c, k]) + 1e-11, (spin, c, k, fd, w2[c, k])


def test_two_particle_loop_converges(m):
    """H4 chain, cc-pVDZ: with the two-particle term the loop still converges
    (to a lower MP2 energy: the rotation now also follows the pair term), and
    its first iteration is the reference loop's first iteration plus the term."""
    mol, S, H, L = system(m, [(1, 0.0, 0.0, 1.4 * k) for k in range(4)], "cc-pvdz", thr=1e-6)
    g = rhf_guess(mol, S, H, L, 2)
    args = (H, L, g["C_a"], g["C_b"], g["P_a"], g["P_b"], g["eps_a"], g["eps_b"], 2, 2, g["e_scf"])
    ref = omp2_ref.cd_omp2(*args)
    r = omp2_ref.cd_omp2(*args, two_particle=True)
    assert ref["converged"] and r["converged"]
    assert r["trace"][0][:2] == ref["trace"][0][:2]
    assert r["trace"][0][2] > ref["trace"][0][2]
    assert r["trace"][-1][2] < 1e-5
    assert r["e_mp2"] < ref["e_mp2"]
    assert m.CholeskyOMP2Config().two_particle_gradient is False


@pytest.mark.parametrize("scale", [1e-3, 0.3, 4.0])
def test_host_expm(m, scale):
    rng = np.random.default_rng(int(scale * 1000))
    A = rng.standard_normal((17, 17))
    K = scale * (A - A.T)
    E = m.linalg_expm(K)
    E_ref = sl.expm(K)
    assert np.abs(E - E_ref).max() < 1e-13 * max(1.0, np.abs(E_ref).max())
    assert np.abs(E.T @ E - np.eye(17)).max() < 1e-13  # exp of antisymmetric is orthogonal
    G = scale * rng.standard_normal((9, 9))
    assert np.abs(m.linalg_expm(G) - sl.expm(G)).max() < 1e-12 * max(1.0, np.abs(sl.expm(G)).max())


def _rot2(B, axis, p, q, th):
    """Rotate rows (axis 1) or columns (axis 2) p, q of B[Q, o, v] by th:
    new p = c p + s q, new q = -s p + c q."""
    c, s = np.cos(th), np.sin(th)
    B = B.copy()
    if axis == 1:
        x, y = B[:, p, :].copy(), B[:, q, :].copy()
        B[:, p, :], B[:, q, :] = c * x + s * y, -s * x + c * y
    else:
        x, y = B[:, :, p].copy(), B[:, :, q].copy()
        B[:, :, p], B[:, :, q] = c * x + s * y, -s * x + c * y
    return B


@pytest.mark.parametrize("th", [0.6, np.pi / 4, -1.1])
def test_opdm_off_diagonal_by_rotation(th):
    """The OPDM's off-diagonal blocks (cholesky_omp2.cc:232-326), pinned
    without restating them: give two occupied (or two virtual) orbitals the
    same energy and rotate them by th.  Denominators are unchanged, the
    amplitudes rotate covariantly, so G' = R G R^T and the rotated diagonal
    c^2 G_pp + s^2 G_qq + 2 c s G_pq must equal dE/deps of the rotated
    orbital, which the pinned UMP2 oracle gives (its dE/deps are pinned by
    central differences and the reference build).  A wrong sign or index in
    G_pq fails the 2 c s G_pq term."""
    from mshqc_amd import synthetic as syn

    (Ba, eoa, eva), (Bb, eob, evb) = syn.ump2_inputs(30, 4, 9, 3, 8, 95)
    eoa[2] = eoa[1]
    eva[5] = eva[3]
    eob[1] = eob[0]
    evb[6] = evb[2]
    G = omp2_ref.opdm(*omp2_ref.t2_blocks(Ba, Bb, eoa, eva, eob, evb))
    c, s = np.cos(th), np.sin(th)
    cases = [  # (spin block rotated, axis, p, q, index of G, index of dE/deps)
        ("a", 1, 1, 2, 0, 2), ("a", 2, 3, 5, 1, 3), ("b", 1, 0, 1, 2, 4), ("b", 2, 2, 6, 3, 5)]
    for spin, axis, p, q, gi, di in cases:
        Ba2 = _rot2(Ba, axis, p, q, th) if spin == "a" else Ba
        Bb2 = _rot2(Bb, axis, p, q, th) if spin == "b" else Bb
        _, grads = oracle.ump2(Ba2, Bb2, eoa, eva, eob, evb, grad=True)
        Gm = G[gi]
        want = c * c * Gm[p, p] + s * s * Gm[q, q] + 2 * c * s * Gm[p, q]
        assert abs(grads[di][p] - want) <= 1e-12 * np.abs(Gm).max(), (spin, axis)
        assert abs(Gm[p, q]) > 1e-6 * np.abs(Gm).max()  # the off-diagonal term is not negligible


def _uhf_energy(H, eri, Ca, Cb, na, nb):
    """E = sum_s tr(P_s H) + 1/2 sum_s tr(P_s (J[P_a + P_b] - K[P_s])) from a
    4-index ERI (the UHF electronic energy, without the oracle's J/K)."""
    Pa, Pb = Ca[:, :na] @ Ca[:, :na].T, Cb[:, :nb] @ Cb[:, :nb].T
    J = np.einsum("mnls,ls->mn", eri, Pa + Pb)
    E = 0.0
    for P in (Pa, Pb):
        K = np.einsum("mlns,ls->mn", eri, P)
        E += np.sum(P * H) + 0.5 * np.sum(P * (J - K))
    return E


def test_rotation_step_is_a_descent_step_with_slope_minus_kappa_w():
    """The loop's step (cholesky_omp2.cc:375-425), pinned by the mathematics
    of its convention: for orbitals C -> C exp(-lambda kappa) the UHF energy
    changes at the rate -sum_ai kappa_ai w_ai, w = 2 F_vo, so w is exactly
    minus the energy gradient in this convention (a sign or factor error in w,
    in kappa's antisymmetry or in exp(-kappa) breaks the equality), and the
    damped, level-shifted step lowers the energy.  The energy comes from a
    brute-force 4-index ERI, not from the oracle's J/K."""
    rng = np.random.default_rng(96)
    n, na, nb, Q = 8, 3, 2, 30
    A = rng.standard_normal((n, n))
    H = 0.5 * (A + A.T) - 2.0 * np.eye(n)
    L = np.empty((Q, n, n))
    for q in range(Q):
        X = 0.05 * rng.standard_normal((n, n))
        L[q] = 0.5 * (X + X.T)
    eri = np.einsum("qmn,qls->mnls", L, L)
    # H's aufbau orbitals (the weak two-electron part keeps its level order),
    # then a small random rotation so w is not zero
    _, U = np.linalg.eigh(H)
    Ca = U @ sl.expm(0.05 * (lambda X: X - X.T)(rng.standard_normal((n, n))))
    Cb = U @ sl.expm(0.05 * (lambda X: X - X.T)(rng.standard_normal((n, n))))
    Pa, Pb = Ca[:, :na] @ Ca[:, :na].T, Cb[:, :nb] @ Cb[:, :nb].T
    J = np.einsum("mnls,ls->mn", eri, Pa + Pb)
    kap, slope = [], 0.0
    for C, P, o in ((Ca, Pa, na), (Cb, Pb, nb)):
        F = C.T @ (H + J - np.einsum("mlns,ls->mn", eri, P)) @ C
        w = 2.0 * F[o:, :o]
        k = omp2_ref.rotation_step(w, F, o)
        assert np.allclose(k, -k.T, atol=0)
        # the descent claim below needs every level-shifted difference positive
        assert np.all(np.diag(F)[o:, None] - np.diag(F)[None, :o] + 0.2 > 0)
        kap.append(k)
        slope -= np.sum(k[o:, :o] * w)
    def cdiff(lam):
        Ep = _uhf_energy(H, eri, Ca @ sl.expm(-lam * kap[0]), Cb @ sl.expm(-lam * kap[1]), na, nb)
        Em = _uhf_energy(H, eri, Ca @ sl.expm(lam * kap[0]), Cb @ sl.expm(lam * kap[1]), na, nb)
        return (Ep - Em) / (2 * lam)

    fd = (4 * cdiff(5e-5) - cdiff(1e-4)) / 3  # Richardson: O(lambda^4)
    assert abs(fd - slope) <= 1e-9 * abs(slope)
    assert slope < 0
    E0 = _uhf_energy(H, eri, Ca, Cb, na, nb)
    E1 = _uhf_energy(H, eri, Ca @ sl.expm(-kap[0]), Cb @ sl.expm(-kap[1]), na, nb)
    assert E1 < E0
