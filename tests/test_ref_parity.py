"""The oracle against the reference's OWN code (oracle/_ref/ref_driver, built
from the reference sources by oracle/ref.mk): pivoted Cholesky vectors,
CD-RMP2 energy and amplitudes, CD-UMP2 energy components, CD-RMP3 energy and
second-order amplitudes, CD-UMP3 energy components.  Skipped when the
reference build is absent (it needs the reference checkout at build time)."""
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
from conftest import ROOT
from oracle import mp3_ref

REF = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
BD = os.path.join(ROOT, "mshqc_amd", "data", "basis")

pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="reference build not present")


def run_ref(eri, thr, Ca, Cb, ea, eb, na, nb):
    n = eri.shape[0]
    with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as f:
        f.write(struct.pack("<iiid", n, na, nb, thr))
        for a in (eri, Ca, Cb, ea, eb):
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        path = f.name
    try:
        out = subprocess.run([REF, path], check=True, capture_output=True, text=True).stdout
    finally:
        os.unlink(path)
    res = {}
    for line in out.splitlines():
        key, *vals = line.split()
        res[key] = np.array([float(x) for x in vals])
    return res


@pytest.fixture(scope="module")
def problem():
    from mshqc_amd import native_module

    m = native_module()
    mol = m.Molecule()
    for k, (x, z) in enumerate([(0.0, 0.0), (0.1, 1.4), (-0.2, 2.9), (0.3, 4.2)]):
        mol.add_atom(1, x, 0.05 * k, z)
    b = m.BasisSet("6-31g", mol, BD)
    eri = m.IntegralEngine(mol, b).compute_eri().numpy()
    n = b.n_basis_functions()
    rng = np.random.default_rng(5)
    Ca, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Cb, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ea = np.sort(np.concatenate([rng.uniform(-1.5, -0.3, 3), rng.uniform(0.1, 2.0, n - 3)]))
    eb = np.sort(np.concatenate([rng.uniform(-1.5, -0.3, 2), rng.uniform(0.1, 2.0, n - 2)]))
    return eri, Ca, Cb, ea, eb


@pytest.mark.parametrize("thr", [1e-3, 1e-6, 1e-10])
def test_oracle_matches_reference_build(problem, thr):
    eri, Ca, Cb, ea, eb = problem
    n = eri.shape[0]
    na, nb = 3, 2
    ref = run_ref(eri, thr, Ca, Cb, ea, eb, na, nb)
    L = oracle.cholesky(eri, thr)
    assert int(ref["nvec"][0]) == L.shape[0]
    assert np.abs(ref["L"].reshape(L.shape) - L).max() < 1e-12
    L3 = L.reshape(-1, n, n)
    # CD-RMP2 (alpha orbitals, n_occ = na)
    B = oracle.transform(L3, Ca[:, :na], Ca[:, na:])
    E, t2 = oracle.rmp2_energy(B, ea[:na], ea[na:], want_t2=True)
    assert abs(ref["rmp2"][0] - E) <= 1e-13 * abs(E)
    assert np.abs(ref["t2"] - t2.ravel()).max() <= 1e-13 * np.abs(t2).max()
    # CD-UMP2 components
    Ba = B
    Bb = oracle.transform(L3, Cb[:, :nb], Cb[:, nb:])
    Eu = oracle.ump2(Ba, Bb, ea[:na], ea[na:], eb[:nb], eb[nb:])
    for got, want in zip(ref["ump2"], Eu):
        assert abs(got - want) <= 1e-13 * max(abs(want), 1e-300) + 1e-17
    # CD-RMP3 from the RMP2 amplitudes (alpha orbitals, n_occ = na)
    Co, Cv = Ca[:, :na], Ca[:, na:]
    E3, t22 = mp3_ref.rmp3(B, oracle.transform(L3, Co, Co), oracle.transform(L3, Cv, Cv), t2, ea[:na], ea[na:])
    assert abs(ref["rmp3"][0] - E3) <= 1e-12 * abs(E3)
    assert np.abs(ref["t22"] - t22.ravel()).max() <= 1e-12 * np.abs(t22).max()
    # CD-UMP3 components (alpha: na occupied of Ca, beta: nb of Cb)
    Cbo, Cbv = Cb[:, :nb], Cb[:, nb:]
    E3u = mp3_ref.ump3(Ba, Bb, oracle.transform(L3, Co, Co), oracle.transform(L3, Cbo, Cbo),
                       oracle.transform(L3, Cv, Cv), oracle.transform(L3, Cbv, Cbv),
                       ea[:na], ea[na:], eb[:nb], eb[nb:])
    for got, want in zip(ref["ump3"], E3u):
        assert abs(got - want) <= 1e-12 * abs(want)
