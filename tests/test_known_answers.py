"""The reference's own published numbers for the path: the Li atom (doublet)
runs of `python test/test_cholesky_sa_caspt3.py` logged in
hasil/hasil_test_FCI.md (lines 7-45, 130-168, 253-291, 379-414): the
basis size, the pivoted CD at 1e-8 of the full ERI matrix (initial maximum
diagonal, vector count) and the CD-UHF energy on those vectors; and the CD
of the cc-pVTZ / cc-pVQZ Li runs in include/mshqc/mcscf/hasilterbaiksaatini.md
(lines 92-121, 179-208), which log no energy.  Here the
build's integral engine and host CholeskyERI::decompose feed the oracle's
CD-UHF restatement (the GPU chain through the drop-in classes is in
test_gpu_dropin.py).  Runs on the CPU.

For the augmented sets the ratio of the largest diagonal to the threshold
(3.4e16 and 1.6e18) exceeds 1/eps of FP64, so the diagonal update's rounding
decides the last pivots and the vector count is pinned to within what that
rounding moves (reference 590 / 1175, here 589 / 1168); the energies are
pinned to the printed digits."""
import os

import numpy as np
import pytest

from conftest import ROOT
from oracle import scf_ref

BD = os.path.join(ROOT, "mshqc_amd", "data", "basis")

# basis, N, vectors, vector-count slack, initial max diagonal, its relative
# tolerance (printed digits), E_UHF, energy tolerance
CASES = [
    ("cc-pcvtz", 43, 394, 0, 8.827692e6, 1e-7, -7.43270280, 6e-9),
    ("aug-cc-pcvtz", 59, 590, 1, 3.4123e8, 5e-5, -7.43270609, 3e-8),
    ("cc-pcvqz", 84, 770, 0, 1.6049e7, 5e-5, -7.43271913, 6e-9),
    ("aug-cc-pcvqz", 109, 1175, 8, 1.5608e10, 5e-5, -7.43271942, 6e-9),
]


@pytest.fixture(scope="module")
def m():
    from mshqc_amd import native_module

    return native_module()


def li(m, basis):
    mol = m.Molecule(0, 2)
    mol.add_atom(3, 0.0, 0.0, 0.0)
    b = m.BasisSet(basis, mol, BD)
    return mol, b, m.IntegralEngine(mol, b)


@pytest.mark.parametrize("basis,N,nvec,slack,maxdiag,dtol,E,etol", CASES)
def test_li_cd_uhf(m, basis, N, nvec, slack, maxdiag, dtol, E, etol):
    mol, b, eng = li(m, basis)
    n = b.n_basis_functions()
    assert n == N
    eri = eng.compute_eri()
    d = np.einsum("ijij->ij", eri.numpy()).max()
    assert abs(d - maxdiag) <= dtol * maxdiag
    chol = m.CholeskyERI(1e-8)
    r = chol.decompose(eri)
    assert r.converged and abs(r.n_vectors - nvec) <= slack
    L = np.array(chol.get_L_vectors()).reshape(-1, n, n)
    H = eng.compute_kinetic() + eng.compute_nuclear()
    u = scf_ref.cd_uhf(eng.compute_overlap(), H, L, 2, 1, mol.nuclear_repulsion_energy(), max_iterations=100)
    assert abs(u["energy"] - E) <= etol, (u["energy"], E)


@pytest.mark.parametrize("basis,N,nvec,maxdiag,dtol", [("cc-pvtz", 30, 316, 8.827692e6, 1e-7),
                                                     ("cc-pvqz", 55, 650, 1.60491e7, 5e-6)])
def test_li_cd_counts(m, basis, N, nvec, maxdiag, dtol):
    mol, b, eng = li(m, basis)
    assert b.n_basis_functions() == N
    eri = eng.compute_eri()
    d = np.einsum("ijij->ij", eri.numpy()).max()
    assert abs(d - maxdiag) <= dtol * maxdiag
    r = m.CholeskyERI(1e-8).decompose(eri)
    assert r.converged and r.n_vectors == nvec
