"""Host-side drop-in pieces that need no GPU: the pivoted Cholesky of
CholeskyERI against the oracle restatement, API conventions and errors."""
import os

import numpy as np
import pytest

import oracle
from conftest import ROOT

BD = os.path.join(ROOT, "mshqc_amd", "data", "basis")


@pytest.fixture(scope="module")
def m():
    from mshqc_amd import native_module

    return native_module()


def h_chain(m, n, spacing=1.4):
    mol = m.Molecule()
    for k in range(n):
        mol.add_atom(1, 0.0, 0.0, spacing * k)
    return mol


def test_cholesky_eri_matches_oracle(m):
    mol = h_chain(m, 4)
    b = m.BasisSet("cc-pvdz", mol, BD)
    eri = m.IntegralEngine(mol, b).compute_eri()
    for thr in (1e-4, 1e-6):
        ce = m.CholeskyERI(thr)
        res = ce.decompose(eri)
        L = np.array(ce.get_L_vectors())
        L_ref = oracle.cholesky(eri.numpy(), thr)
        assert res.n_vectors == ce.n_vectors() == L_ref.shape[0]
        assert np.abs(L - L_ref).max() < 1e-12
        n = b.n_basis_functions()
        full = eri.numpy().reshape(n * n, n * n)
        d = np.diag(full - L.T @ L)
        assert d.max() < thr  # reconstruction on the diagonal
        assert abs(ce.reconstruct(0, 1, 2, 3) - L[:, 0 * n + 1] @ L[:, 2 * n + 3]) < 1e-14


def test_cholesky_eri_edge_cases(m):
    with pytest.raises(ValueError):
        m.CholeskyERI(0.0)
    mol = h_chain(m, 2)
    eri = m.IntegralEngine(mol, m.BasisSet("sto-3g", mol, BD)).compute_eri()
    ce = m.CholeskyERI(1e9)  # above every diagonal: no vectors
    assert ce.decompose(eri).n_vectors == 0
    ce2 = m.CholeskyERI()
    L = np.random.default_rng(0).standard_normal((3, 2, 2))
    ce2.set_vectors(L)
    assert ce2.n_vectors() == 3 and ce2.n_basis() == 2
    ce2.add_vector(np.ones(4))
    assert ce2.n_vectors() == 4


def test_rhf_rejects_odd_electrons(m):
    mol = h_chain(m, 3)
    b = m.BasisSet("sto-3g", mol, BD)
    eng = m.IntegralEngine(mol, b)
    with pytest.raises(RuntimeError, match="closed-shell"):
        m.CholeskyRHF(mol, b, eng, m.CholeskyRHFConfig())


def _eigh_cases():
    rng = np.random.default_rng(1)
    out = []
    for n in (1, 2, 3, 30, 257):
        A = rng.standard_normal((n, n))
        out.append(("random", A + A.T))
    Q, _ = np.linalg.qr(rng.standard_normal((50, 50)))
    out.append(("degenerate", Q @ np.diag(np.repeat([1.0, 2.0, 3.0, -1.0, 5.0], 10)) @ Q.T))
    out.append(("graded", Q @ np.diag(10.0 ** np.linspace(-8, 8, 50)) @ Q.T))
    out.append(("tridiagonal", np.diag(np.full(40, 2.0)) + np.diag(np.full(39, -1.0), 1)
                + np.diag(np.full(39, -1.0), -1)))
    out.append(("zero", np.zeros((6, 6))))
    B = np.zeros((10, 10))
    B[:5, :5] = 1.0
    B[5:, 5:] = 2.0 * np.eye(5)
    out.append(("block", B))
    return out


@pytest.mark.parametrize("kind,A", _eigh_cases(), ids=lambda x: x if isinstance(x, str) else None)
def test_eigh_matches_numpy(m, kind, A):
    """Householder + implicit QL: eigenvalues equal numpy's (LAPACK) to
    1e-13 of the spectral radius, orthonormal eigenvectors, A = V diag(w) V^T,
    ascending order; degenerate, graded, reducible and zero matrices."""
    w, V = m.linalg_eigh(A)
    sc = max(1.0, np.abs(np.linalg.eigvalsh(A)).max())
    assert np.all(np.diff(w) >= 0)
    assert np.abs(w - np.linalg.eigvalsh(A)).max() <= 1e-13 * sc
    assert np.abs(V.T @ V - np.eye(len(w))).max() <= 1e-12
    assert np.abs(V @ np.diag(w) @ V.T - A).max() <= 1e-12 * sc


def test_config_defaults_follow_reference(m):
    c = m.CholeskyRHFConfig()
    assert c.cholesky_threshold == 1e-6 and c.screen_exchange and c.max_iterations == 100
    assert c.energy_threshold == 1e-8 and c.diis_max_vectors == 8
    u = m.CholeskyUHFConfig()
    assert u.diis_max_vectors == 6 and u.max_iterations == 100
    r = m.CholeskyRMP2Config()
    assert r.precision == 0 and r.cholesky_threshold == 1e-6


def test_solve_full_pivot(m):
    """Host LU (DIIS) and the oracle's restatement: numpy's answer on regular
    systems; the singular DIIS system of vanishing error vectors keeps
    sum(c) = 1 with the free coefficients at zero."""
    from oracle.scf_ref import fullpiv_solve

    rng = np.random.default_rng(3)
    for n in (1, 2, 5, 9):
        A = rng.standard_normal((n, n))
        b = rng.standard_normal(n)
        x_ref = np.linalg.solve(A, b)
        assert np.allclose(m.linalg_solve(A, b), x_ref, rtol=1e-10, atol=1e-12)
        assert np.allclose(fullpiv_solve(A, b), x_ref, rtol=1e-10, atol=1e-12)
    for nv in (2, 3, 5):
        A = np.zeros((nv + 1, nv + 1))
        A[:nv, nv] = A[nv, :nv] = -1.0
        b = np.zeros(nv + 1)
        b[nv] = -1.0
        A[:nv, :nv] = 1e-40 * np.eye(nv)  # below the rank threshold
        for x in (m.linalg_solve(A, b), fullpiv_solve(A, b)):
            assert x[:nv].sum() == 1.0 and np.count_nonzero(x[:nv]) == 1 and x[nv] == 0.0
        assert np.array_equal(m.linalg_solve(A, b), fullpiv_solve(A, b))
    # a DIIS system of well-conditioned error vectors: both agree to rounding
    E = rng.standard_normal((4, 30))
    A = np.zeros((5, 5))
    A[:4, :4] = E @ E.T
    A[:4, 4] = A[4, :4] = -1.0
    b = np.zeros(5)
    b[4] = -1.0
    assert np.allclose(m.linalg_solve(A, b), fullpiv_solve(A, b), rtol=1e-12, atol=1e-13)
    assert np.allclose(fullpiv_solve(A, b), np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
