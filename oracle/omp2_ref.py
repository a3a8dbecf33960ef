"""numpy restatement of CD-OMP2 (src/mp2/cholesky_omp2.cc) -- TEST
INFRASTRUCTURE ONLY (tiny sizes).

* ``t2_blocks``  cholesky_omp2.cc:94-183  antisymmetrised t_aa, t_bb and t_ab
  from B^Q (skip |D| < 1e-12)
* ``energy``     cholesky_omp2.cc:189-230  sum_{i<j,a<b} t^2 D (same spin) +
  sum t_ab^2 D (opposite spin)
* ``opdm``       cholesky_omp2.cc:232-326  G_oo (negative) and G_vv per spin.
  Its diagonals are the orbital-energy derivatives of the MP2 energy, which
  is how tests pin the backward kernel's dE/deps against the repository's
  own OMP2 orbital-gradient terms.
* ``cd_omp2``    cholesky_omp2.cc:331-513  the orbital loop: P + C G C^T per
  spin, generalised Fock H + J - K_s in the MO basis, w = 2 F_vo, damped
  level-shifted rotation C exp(-kappa) (scipy's expm); denominators keep
  the SCF orbital energies; on exit the total is the previous iteration's.
  Pinned by: iteration 1 reproducing the UMP2/RMP2 energies of the reference
  build's own code (via the pinned pair oracle), a one-electron system
  stopping at iteration 1 with the SCF energy (w = 2 F_vo is the SCF
  orbital gradient), and orthonormal rotated orbitals (tests/test_omp2.py).
  With ``two_particle=True`` (not in the reference) each spin's w also gets
  ``two_particle_grad``.
* ``rotation_step``  the damped, level-shifted kappa of one iteration
  (cholesky_omp2.cc:375-425).  Pinned by: -sum kappa w is the first-order
  change of the UHF energy along C exp(-lambda kappa), i.e. w = 2 F_vo is
  minus the energy's gradient in this convention, and the step lowers it
  (tests/test_omp2.py, with a brute-force 4-index energy).
* ``opdm`` off-diagonal blocks are pinned by rotating two orbitals of equal
  energy: the rotated diagonal c^2 G_ii + s^2 G_jj + 2 c s G_ij must equal
  the pinned dE/deps of the rotated inputs (tests/test_omp2.py).
* ``two_particle_grad``  the two-particle term the reference's w = 2 F_vo
  leaves out (cholesky_omp2.cc:331-369 build F from the OPDM only):
  w2(c, k) = dE_MP2/dkappa_ck at fixed denominators, from dE/dB and the
  oo / vv blocks.  Pinned by central differences of the MP2 energy under
  the orbital rotation C -> C expm(kappa) (tests/test_omp2.py).
"""
from __future__ import annotations

import numpy as np


def _g(Bl, Br):
    # g[i, j, a, b] = sum_Q Bl[Q, i, a] Br[Q, j, b]
    return np.einsum("qia,qjb->ijab", Bl, Br)


def _denom(eo_l, ev_l, eo_r, ev_r):
    return eo_l[:, None, None, None] + eo_r[None, :, None, None] - ev_l[None, None, :, None] - ev_r[None, None, None, :]


def t2_blocks(Ba, Bb, eoa, eva, eob, evb):
    out = []
    for B, eo, ev in ((Ba, eoa, eva), (Bb, eob, evb)):
        g = _g(B, B)
        D = _denom(eo, ev, eo, ev)
        A = g - g.transpose(0, 1, 3, 2)
        t = np.where(np.abs(D) < 1e-12, 0.0, A / np.where(np.abs(D) < 1e-12, 1.0, D))
        o, v = B.shape[1], B.shape[2]
        # the reference fills only i<j, a<b and mirrors with signs; the diagonal stays 0
        mask = np.zeros((o, o, v, v), bool)
        iu = np.triu(np.ones((o, o), bool), 1)
        au = np.triu(np.ones((v, v), bool), 1)
        mask = iu[:, :, None, None] & au[None, None, :, :]
        tt = np.where(mask, t, 0.0)
        tt = tt - tt.transpose(1, 0, 2, 3)
        tt = tt - tt.transpose(0, 1, 3, 2)
        out.append(tt)
    g = _g(Ba, Bb)
    D = _denom(eoa, eva, eob, evb)
    tab = np.where(np.abs(D) < 1e-12, 0.0, g / np.where(np.abs(D) < 1e-12, 1.0, D))
    return out[0], out[1], tab


def energy(taa, tbb, tab, eoa, eva, eob, evb):
    E = 0.0
    for t, eo, ev in ((taa, eoa, eva), (tbb, eob, evb)):
        D = _denom(eo, ev, eo, ev)
        o, v = len(eo), len(ev)
        m = np.triu(np.ones((o, o), bool), 1)[:, :, None, None] & np.triu(np.ones((v, v), bool), 1)[None, None]
        E += np.sum(np.where(m, t * t * D, 0.0))
    E += np.sum(tab * tab * _denom(eoa, eva, eob, evb))
    return E


def opdm(taa, tbb, tab):
    """(G_oo_a, G_vv_a, G_oo_b, G_vv_b) of cholesky_omp2.cc:232-326."""
    Goo_a = -(0.5 * np.einsum("ikab,jkab->ij", taa, taa) + np.einsum("ikab,jkab->ij", tab, tab))
    Gvv_a = 0.5 * np.einsum("ijac,ijbc->ab", taa, taa) + np.einsum("ijac,ijbc->ab", tab, tab)
    Goo_b = -(0.5 * np.einsum("ikab,jkab->ij", tbb, tbb) + np.einsum("kiab,kjab->ij", tab, tab))
    Gvv_b = 0.5 * np.einsum("ijac,ijbc->ab", tbb, tbb) + np.einsum("ijca,ijcb->ab", tab, tab)
    return Goo_a, Gvv_a, Goo_b, Gvv_b


def two_particle_grad(dB, Loo, Lvv):
    """w2[c, k] = sum_Q [sum_a Lvv[Q, c, a] dB[Q, k, a] - sum_i dB[Q, i, c] Loo[Q, i, k]]:
    B^Q_ia = C_i^T L^Q C_a changes by sum_c kappa_ci Lvv^Q(c, a) - sum_k kappa_ak
    Loo^Q(i, k) when occupied k gains kappa_ck C_c and virtual c loses
    kappa_ck C_k, and the chain rule through dE/dB gives w2."""
    return np.einsum("qca,qka->ck", Lvv, dB) - np.einsum("qic,qik->ck", dB, Loo)


def rotation_step(w, F, o):
    """The reference's orbital step (cholesky_omp2.cc:375-425): the n x n
    antisymmetric kappa with kappa[o + a, i] = 0.5 w[a, i] / d and
    kappa[i, o + a] = -kappa[o + a, i], d = F[o+a, o+a] - F[i, i] + 0.2 with
    |d| raised to at least 1e-2 (sign kept); orbitals then go to C exp(-kappa).
    w is [v, o], F the MO-basis Fock matrix."""
    n = F.shape[0]
    kappa = np.zeros((n, n))
    for a in range(n - o):
        for i in range(o):
            diff = F[o + a, o + a] - F[i, i] + 0.2
            if abs(diff) < 1e-2:
                diff = 1e-2 if diff >= 0 else -1e-2
            val = (w[a, i] / diff) * 0.5
            kappa[o + a, i] = val
            kappa[i, o + a] = -val
    return kappa


def cd_omp2(H, L, C_a, C_b, P_a, P_b, eps_a, eps_b, na, nb, e_scf, max_iterations=50,
            energy_threshold=1e-6, gradient_threshold=1e-5, two_particle=False):
    """The CD-OMP2 loop of cholesky_omp2.cc:427-513 with numpy and the C
    oracle's transform and J/K.  Returns dict(energy_total, e_mp2, e_ss, e_os,
    C_a, C_b, iterations, converged, trace) where trace lists (E_total,
    E_mp2, g_norm) per iteration."""
    import scipy.linalg as sl

    from . import jk as _jk, transform as _transform, ump2 as _ump2

    n = H.shape[0]
    va, vb = n - na, n - nb
    eoa, eva, eob, evb = eps_a[:na], eps_a[na:], eps_b[:nb], eps_b[nb:]
    C_a, C_b, P_a, P_b = C_a.copy(), C_b.copy(), P_a.copy(), P_b.copy()
    Q = L.shape[0]

    def B_of(C, o):
        if o == 0 or o == n:
            return np.zeros((Q, o, n - o))
        return _transform(L, C[:, :o], C[:, o:])

    def rotate(w, F, o, C):
        # :375-425: damped, level-shifted step, C <- C exp(-kappa)
        C = C @ sl.expm(-rotation_step(w, F, o))
        return C, C[:, :o] @ C[:, :o].T

    e_prev = e_scf
    e_mp2 = e_ss = e_os = 0.0
    converged = False
    trace = []
    it = 0
    while it < max_iterations:
        it += 1
        Ba, Bb = B_of(C_a, na), B_of(C_b, nb)
        taa, tbb, tab = t2_blocks(Ba, Bb, eoa, eva, eob, evb)
        e_mp2 = energy(taa, tbb, tab, eoa, eva, eob, evb)
        e_os = float(np.sum(tab * tab * _denom(eoa, eva, eob, evb)))
        e_ss = e_mp2 - e_os
        e_curr = e_scf + e_mp2
        Goo_a, Gvv_a, Goo_b, Gvv_b = opdm(taa, tbb, tab)
        Pt = []
        for C, P, Goo, Gvv, o in ((C_a, P_a, Goo_a, Gvv_a, na), (C_b, P_b, Goo_b, Gvv_b, nb)):
            G = np.zeros((n, n))
            G[:o, :o] = Goo
            G[o:, o:] = Gvv
            Pt.append(P + C @ G @ C.T)
        J, K = _jk(L, np.stack(Pt))
        F_a = C_a.T @ (H + J - K[0]) @ C_a
        F_b = C_b.T @ (H + J - K[1]) @ C_b
        w_a = 2.0 * F_a[na:, :na]
        w_b = 2.0 * F_b[nb:, :nb]
        if two_particle:
            _, (dBa, dBb, *_) = _ump2(Ba, Bb, eoa, eva, eob, evb, grad=True)
            w_a = w_a + two_particle_grad(dBa, _transform(L, C_a[:, :na], C_a[:, :na]),
                                          _transform(L, C_a[:, na:], C_a[:, na:]))
            w_b = w_b + two_particle_grad(dBb, _transform(L, C_b[:, :nb], C_b[:, :nb]),
                                          _transform(L, C_b[:, nb:], C_b[:, nb:]))
        g_norm = float(np.sqrt(np.sum(w_a * w_a) + np.sum(w_b * w_b)))
        trace.append((e_curr, e_mp2, g_norm))
        if g_norm < gradient_threshold and abs(e_curr - e_prev) < energy_threshold:
            converged = True
            break
        C_a, P_a = rotate(w_a, F_a, na, C_a)
        if na != nb:
            C_b, P_b = rotate(w_b, F_b, nb, C_b)
        else:
            C_b, P_b = C_a, P_a
        e_prev = e_curr
    return dict(energy_total=e_prev, e_mp2=e_mp2, e_ss=e_ss, e_os=e_os, C_a=C_a, C_b=C_b,
                iterations=it, converged=converged, trace=trace)
