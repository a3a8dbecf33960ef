"""CD-ROHF: the oracle's restatement of the reference driver
(src/scf/cholesky_rohf.cc) pinned by cases whose answers the mathematics
fixes, and the occupied-factor exchange it relies on.  Runs on the CPU with
the build's own integral engine and host-side Cholesky decomposition."""
import os

import numpy as np
import pytest
import scipy.linalg as sl

import oracle
from conftest import ROOT
from mshqc_amd import synthetic as syn
from oracle import scf_ref

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


def test_one_electron_is_exact(m):
    """H atom (1 alpha, 0 beta): J - K annihilates the occupied orbital, so
    the energy is the lowest eigenvalue of H (cc-pVDZ: -0.49927840)."""
    mol, S, H, L = system(m, [(1, 0.0, 0.0, 0.0)], "cc-pvdz", mult=2)
    r = scf_ref.cd_rohf(S, H, L, 1, 0, mol.nuclear_repulsion_energy())
    e0 = sl.eigh(H, S, eigvals_only=True)[0]
    assert r["converged"]
    assert abs(r["energy"] - e0) < 1e-12
    assert abs(r["energy"] - (-0.49927840)) < 2e-8
    assert np.all(r["Pb"] == 0.0)


def test_closed_shell_equals_rhf(m):
    """n_a = n_b: F_a = F_b = H + J[2P] - K[P], the RHF Fock matrix, so the
    converged energy is the CD-RHF one."""
    mol, S, H, L = system(m, [(1, 0.0, 0.0, 1.4 * k) for k in range(4)], "cc-pvdz")
    e_nuc = mol.nuclear_repulsion_energy()
    ro = scf_ref.cd_rohf(S, H, L, 2, 2, e_nuc, energy_threshold=1e-12, max_iterations=300)
    rh = scf_ref.cd_rhf(S, H, L, 2, e_nuc, energy_threshold=1e-12, max_iterations=300)
    assert ro["converged"]
    assert abs(ro["energy"] - rh["energy"]) < 1e-9
    assert np.abs(ro["Pa"] - ro["Pb"]).max() == 0.0


def test_open_shell_bounded_by_uhf(m):
    """Li (2 alpha, 1 beta), 6-31G: a restricted determinant is a special
    UHF determinant, so E(ROHF) >= E(UHF); both near -7.4315."""
    mol, S, H, L = system(m, [(3, 0.0, 0.0, 0.0)], "6-31g", mult=2)
    e_nuc = mol.nuclear_repulsion_energy()
    ro = scf_ref.cd_rohf(S, H, L, 2, 1, e_nuc, energy_threshold=1e-12, max_iterations=300)
    uh = scf_ref.cd_uhf(S, H, L, 2, 1, e_nuc, energy_threshold=1e-12, max_iterations=300)
    print(f"Li 6-31G: ROHF {ro['energy']!r} ({ro['iterations']} it)  UHF {uh['energy']!r}")
    assert ro["converged"]
    assert ro["energy"] >= uh["energy"] - 1e-10
    assert ro["energy"] - uh["energy"] < 2e-3
    # shared spatial orbitals: P_b is the leading block of P_a
    C = ro["C"]
    assert np.abs(ro["Pa"] - ro["Pb"] - np.outer(C[:, 1], C[:, 1])).max() < 1e-14


@pytest.mark.parametrize("na,nb", [(7, 4), (5, 0), (6, 6)])
def test_occupied_factor_exchange(na, nb):
    """sum_Q (L^Q C_s)(L^Q C_s)^T = sum_Q L^Q P_s L^Q with P_s = C_s C_s^T (the
    exchange form of cholesky_rohf.cc:125-133), and J of Pa + Pb."""
    rng = np.random.default_rng(na * 10 + nb)
    Q, n = 17, 23
    L = syn.gen_L(Q, n, rng)
    C = syn.gen_Cocc(n, n, rng)
    X = oracle.half_transform(L, C[:, :max(na, nb)])
    Ka = np.einsum("qmi,qni->mn", X[:, :, :na], X[:, :, :na])
    Kb = np.einsum("qmi,qni->mn", X[:, :, :nb], X[:, :, :nb])
    Pa, Pb = C[:, :na] @ C[:, :na].T, C[:, :nb] @ C[:, :nb].T
    J, K = oracle.jk(L, np.stack([Pa, Pb]))
    assert np.abs(Ka - K[0]).max() < 1e-12
    assert np.abs(Kb - K[1]).max() < 1e-12
    J_ref = np.einsum("qkl,kl,qmn->mn", L, Pa + Pb, L)
    assert np.abs(J - J_ref).max() < 1e-12
