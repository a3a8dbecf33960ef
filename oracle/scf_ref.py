"""numpy restatement of the CD-RHF / CD-UHF / CD-ROHF drivers -- TEST INFRASTRUCTURE ONLY.

Used by tests to check the drop-in SCF classes, whose J/K go through the HIP
library.  Control flow follows the reference drivers step by step:

* ``cd_rhf``: src/scf/cholesky_rhf.cc:157-254 -- core-Hamiltonian guess
  (:80-88), F = H + J - K/2 (:99-137, with screen_exchange), the DIIS block
  that never engages (:188-196), 0.7/0.3 density damping (:206-210),
  E = 1/2 tr P (H + F) + E_nuc (:142-144), stop on |dE| < thr after iter 1.
* ``cd_uhf``: src/scf/cholesky_uhf.cc:121-159 -- guess with -0.5 on the
  diagonal (:68-75), DIIS stored every iteration and used from iteration 3,
  damping from iteration 2, stop on |dE| < thr.
* ``cd_rohf``: src/scf/cholesky_rohf.cc:184-276 -- one core-Hamiltonian
  guess for both spins (:77-83), F_s = H + J[Pa + Pb] - K[P_s] (:89-147,
  here through the general-density J/K, not the occupied-factor form the
  library uses), electronic E = 1/2 sum P_s * (H + F_s) (:159-165), per-spin
  DIIS only while ||FPS - SPF|| < diis_threshold (:214-226), orbitals of
  (n_a F_a + n_b F_b)/(n_a + n_b) shared by both spins (:228-236), no damping,
  stop on |dE| < thr after iteration 1.

Parity: pinned by the textbook RHF energies of H2/STO-3G and the H-atom
cc-pVDZ energy (tests/test_integrals.py) and by CD -> conventional RHF as
the threshold goes to 0.  ``cd_rohf`` is pinned by the exact one-electron
case (H atom: the lowest eigenvalue of H), by n_a = n_b reducing to the
converged CD-RHF energy, and by the UHF lower bound for an open shell
(tests/test_rohf.py).
"""
from __future__ import annotations

import numpy as np

from . import jk as _jk


def _inv_sqrt(S):
    w, U = np.linalg.eigh(S)
    return (U / np.sqrt(w)) @ U.T


def _solve(F, X):
    e, U = np.linalg.eigh(X.T @ F @ X)
    return e, X @ U


def fullpiv_solve(A, b):
    """A x = b as Eigen's FullPivLU::solve does it (the reference's DIIS,
    src/scf/diis.cc:73): pivot = first largest |a_ij| scanning column by
    column; elimination stops at an all-zero corner; pivots at or below
    n * eps * (largest pivot) count as zero and their unknowns are 0."""
    A = np.array(A, dtype=np.float64)
    b = np.array(b, dtype=np.float64)
    n = A.shape[0]
    perm = np.arange(n)
    nz, maxpiv = n, 0.0
    for k in range(n):
        corner = np.abs(A[k:, k:])
        flat = int(np.argmax(corner.T.ravel()))  # column-major, first maximum
        pc, pr = divmod(flat, n - k)
        best = corner[pr, pc]
        if best == 0.0:
            nz = k
            break
        maxpiv = max(maxpiv, best)
        pr += k
        pc += k
        A[[k, pr]] = A[[pr, k]]
        b[[k, pr]] = b[[pr, k]]
        A[:, [k, pc]] = A[:, [pc, k]]
        perm[[k, pc]] = perm[[pc, k]]
        for i in range(k + 1, n):
            f = A[i, k] / A[k, k]
            A[i, k + 1:] -= f * A[k, k + 1:]
            A[i, k] = f
            b[i] -= f * b[k]
    thr = maxpiv * np.finfo(np.float64).eps * n
    rank = int(sum(abs(A[k, k]) > thr for k in range(nz)))
    y = np.zeros(n)
    for i in range(rank - 1, -1, -1):
        y[i] = (b[i] - A[i, i + 1:rank] @ y[i + 1:rank]) / A[i, i]
    x = np.zeros(n)
    x[perm] = y
    return x


class _DIIS:
    """Pulay DIIS with the bookkeeping of src/scf/diis.cc:36-108."""

    def __init__(self, nmax):
        self.nmax, self.F, self.E = nmax, [], []

    def add(self, F, err):
        self.F.append(F)
        self.E.append(err)
        if len(self.F) > self.nmax:
            self.F.pop(0)
            self.E.pop(0)

    def extrapolate(self):
        n = len(self.F)
        A = np.zeros((n + 1, n + 1))
        for i in range(n):
            for j in range(n):
                A[i, j] = np.sum(self.E[i] * self.E[j])
        A[:n, n] = A[n, :n] = -1.0
        b = np.zeros(n + 1)
        b[n] = -1.0
        c = fullpiv_solve(A, b)[:n]
        return sum(ci * Fi for ci, Fi in zip(c, self.F))


def cd_rhf(S, H, L, n_occ, e_nuc, max_iterations=100, energy_threshold=1e-8, screen=True):
    """Returns dict(energy, C, eps, P, F, iterations, converged)."""
    X = _inv_sqrt(S)
    eps, C = _solve(H, X)
    P = 2.0 * C[:, :n_occ] @ C[:, :n_occ].T
    energy = 0.0
    converged = False
    F = H
    it = 0
    for it in range(1, max_iterations + 1):
        e_old = energy
        P_prev = P
        J, K = _jk(L, P, screen=screen)
        F = H + J - 0.5 * K
        eps, C = _solve(F, X)
        P_new = 2.0 * C[:, :n_occ] @ C[:, :n_occ].T
        P = 0.7 * P_new + 0.3 * P_prev
        energy = 0.5 * np.sum(P * (H + F).T) + e_nuc
        de = energy - e_old if it > 1 else energy
        if abs(de) < energy_threshold and it > 1:
            converged = True
            break
    return dict(energy=energy, C=C, eps=eps, P=P, F=F, iterations=it, converged=converged)


def cd_uhf(S, H, L, na, nb, e_nuc, max_iterations=100, energy_threshold=1e-8, diis_max=6):
    X = _inv_sqrt(S)
    Hb = H - 0.5 * np.eye(H.shape[0])
    ea, Ca = _solve(Hb, X)
    eb, Cb = _solve(Hb, X)
    Pa = Ca[:, :na] @ Ca[:, :na].T
    Pb = Cb[:, :nb] @ Cb[:, :nb].T
    da, db = _DIIS(diis_max), _DIIS(diis_max)
    energy = 0.0
    it = 0
    for it in range(1, max_iterations + 1):
        e_old = energy
        J, K = _jk(L, np.stack([Pa, Pb]))
        Fa = H + J - K[0]
        Fb = H + J - K[1]
        da.add(Fa, Fa @ Pa @ S - S @ Pa @ Fa)
        db.add(Fb, Fb @ Pb @ S - S @ Pb @ Fb)
        if it > 2:
            Fa, Fb = da.extrapolate(), db.extrapolate()
        ea, Ca = _solve(Fa, X)
        eb, Cb = _solve(Fb, X)
        Pa_new = Ca[:, :na] @ Ca[:, :na].T
        Pb_new = Cb[:, :nb] @ Cb[:, :nb].T
        if it > 1:
            Pa, Pb = 0.7 * Pa_new + 0.3 * Pa, 0.7 * Pb_new + 0.3 * Pb
        else:
            Pa, Pb = Pa_new, Pb_new
        energy = 0.5 * (np.sum(Pa * (H + Fa).T) + np.sum(Pb * (H + Fb).T))
        if abs(energy - e_old) < energy_threshold:
            break
    return dict(energy=energy + e_nuc, Ca=Ca, Cb=Cb, ea=ea, eb=eb, Pa=Pa, Pb=Pb, iterations=it)


def cd_rohf(S, H, L, na, nb, e_nuc, max_iterations=100, energy_threshold=1e-8, diis_threshold=1e-2,
            diis_max=8):
    """Returns dict(energy (total), e_elec, C, eps, Pa, Pb, iterations, converged)."""
    X = _inv_sqrt(S)
    eps, C = _solve(H, X)

    def dens(m):
        return C[:, :m] @ C[:, :m].T

    Pa, Pb = dens(na), dens(nb)
    da, db = _DIIS(diis_max), _DIIS(diis_max)
    energy = 0.0
    converged = False
    it = 0
    while it < max_iterations:
        it += 1
        e_old = energy
        J, K = _jk(L, np.stack([Pa, Pb]))
        Fa = H + J - K[0]
        Fb = H + J - K[1]
        energy = 0.5 * np.sum(Pa * (H + Fa)) + 0.5 * np.sum(Pb * (H + Fb))
        for d, F, P, side in ((da, Fa, Pa, 0), (db, Fb, Pb, 1)):
            err = F @ P @ S - S @ P @ F
            if np.linalg.norm(err) < diis_threshold:
                d.add(F, err)
                if len(d.F) >= 2:
                    if side == 0:
                        Fa = d.extrapolate()
                    else:
                        Fb = d.extrapolate()
        eps, C = _solve((na * Fa + nb * Fb) / (na + nb), X)
        Pa, Pb = dens(na), dens(nb)
        if abs(energy - e_old) < energy_threshold and it > 1:
            converged = True
            break
    return dict(energy=energy + e_nuc, e_elec=energy, C=C, eps=eps, Pa=Pa, Pb=Pb, iterations=it,
                converged=converged)
