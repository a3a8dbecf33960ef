"""numpy restatement of the CD-OMP2 amplitude, energy and one-particle
density pieces -- TEST INFRASTRUCTURE ONLY (tiny sizes).

* ``t2_blocks``  cholesky_omp2.cc:94-183  antisymmetrised t_aa, t_bb and t_ab
  from B^Q (skip |D| < 1e-12)
* ``energy``     cholesky_omp2.cc:189-230  sum_{i<j,a<b} t^2 D (same spin) +
  sum t_ab^2 D (opposite spin)
* ``opdm``       cholesky_omp2.cc:232-326  G_oo (negative) and G_vv per spin

Its diagonals are the orbital-energy derivatives of the MP2 energy, which
is how tests pin the backward kernel's dE/deps against the repository's own
OMP2 orbital-gradient terms.
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
