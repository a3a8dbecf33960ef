"""Pivoted Cholesky of the ERI matrix on the GPU against the FP64 oracle
(whose vectors match the reference's compiled CholeskyERI::decompose) and the
host decomposition: the same pivots and bitwise the same vectors."""
import os

import numpy as np
import pytest

import oracle
from conftest import ROOT, to_dev, to_np

pytestmark = pytest.mark.gpu
BD = os.path.join(ROOT, "mshqc_amd", "data", "basis")


@pytest.fixture(scope="module")
def m(ctx):
    from mshqc_amd import native_module

    return native_module()


def _eri(m, atoms, basis):
    mol = m.Molecule()
    for Z, x, y, z in atoms:
        mol.add_atom(Z, x, y, z)
    b = m.BasisSet(basis, mol, BD)
    eng = m.IntegralEngine(mol, b)
    e = eng.compute_eri()
    n = b.n_basis_functions()
    return e, e.numpy().reshape(n * n, n * n)


@pytest.mark.parametrize("natoms,thr", [(4, 1e-6), (4, 1e-10), (10, 1e-6), (10, 1e-8)])
def test_cholesky_matches_oracle_bitwise(ctx, m, natoms, thr):
    e, M = _eri(m, [(1, 0.0, 0.0, 1.4 * k) for k in range(natoms)], "cc-pvdz")
    print(f"cholesky H{natoms} n2={M.shape[0]} thr={thr}", flush=True)
    n = int(round(M.shape[0] ** 0.5))
    L_ref = oracle.cholesky(M.reshape(n, n, n, n), thr)
    L, conv, res = ctx.cholesky(*to_dev(M), thr)
    L = to_np(L)
    assert conv and res < thr
    assert L.shape == L_ref.shape
    assert np.array_equal(L, L_ref)
    # the drop-in class: device path equals the host path
    c_gpu, c_host = m.CholeskyERI(thr), m.CholeskyERI(thr)
    r_gpu, r_host = c_gpu.decompose_gpu(e), c_host.decompose(e)
    assert r_gpu.n_vectors == r_host.n_vectors == L.shape[0] and r_gpu.converged
    assert np.array_equal(np.array(c_gpu.get_L_vectors()), np.array(c_host.get_L_vectors()))


def test_cholesky_edges(ctx, m):
    e, M = _eri(m, [(1, 0.0, 0.0, 0.0), (1, 0.0, 0.0, 1.4)], "cc-pvdz")
    dM = to_dev(M)[0]
    L, conv, res = ctx.cholesky(dM, 1e3)  # above every diagonal: no vectors
    assert L.shape[0] == 0 and conv and res < 1e3
    full, conv, _ = ctx.cholesky(dM, 1e-9)
    assert conv
    part, conv, res = ctx.cholesky(dM, 1e-9, max_vectors=5)  # stopped by the bound
    assert part.shape[0] == 5 and not conv and res >= 1e-9
    assert np.array_equal(to_np(part), to_np(full[:5]))
    # exact reconstruction at a tiny threshold
    assert np.abs(to_np(full).T @ to_np(full) - M).max() < 1e-9
