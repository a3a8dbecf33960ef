"""numpy restatement of CD-RMP3 (src/mp3/cholesky_rmp3.cc:79-200) and CD-UMP3
(src/mp3/cholesky_ump3.cc:160-637) -- TEST INFRASTRUCTURE ONLY (tiny sizes).

Inputs are the MO Cholesky blocks B^Q_ia, Loo^Q_ki, Lvv^Q_ac (each
C_left^T L^Q C_right, :33-77) and the first-order amplitudes t1 (the CD-RMP2
t2, t_ij^ab = (ia|jb)/D).  Second-order amplitudes (:113-163):
  D t2_2(ijab) = 1/2 sum_cd [(ac|bd) - (ad|bc)] t1(ijcd)
               + 1/2 sum_kl [(ki|lj) - (kj|li)] t1(klab)
               - sum_kc [2(jb|kc) - (kb|jc)] t1(ikac)
               - sum_kc [2(ia|kc) - (ka|ic)] t1(jkbc)
(|D| < 1e-12 leaves t2_2 = 0) and E3 = sum_ijab [2(ia|jb) - (ib|ja)] t2_2
(:166-178).  This is the reference's own formula, not the textbook
closed-shell MP3, so it is pinned against the reference's compiled
CholeskyRMP3 (oracle/_ref, tests/test_ref_parity.py).

``ump3`` restates CholeskyUMP3's energies with the reference's own index
conventions, including its reading of column-major o x v blocks as
row-major (``scramble``); it is pinned the same way.
"""
from __future__ import annotations

import numpy as np


def rmp3(B, Loo, Lvv, t1, eps_o, eps_v):
    """Returns (E3, t2_2[o,o,v,v])."""
    V = np.einsum("qac,qbd->acbd", Lvv, Lvv)          # (ac|bd)
    num = 0.5 * (np.einsum("acbd,ijcd->ijab", V, t1) - np.einsum("adbc,ijcd->ijab", V, t1))
    W = np.einsum("qki,qlj->kilj", Loo, Loo)          # (ki|lj)
    num += 0.5 * (np.einsum("kilj,klab->ijab", W, t1) - np.einsum("kjli,klab->ijab", W, t1))
    G = np.einsum("qia,qjb->iajb", B, B)              # (ia|jb)
    Y = 2.0 * G - G.transpose(2, 1, 0, 3)             # Y[j,b,k,c] = 2(jb|kc) - (kb|jc)
    num -= np.einsum("jbkc,ikac->ijab", Y, t1)
    num -= np.einsum("iakc,jkbc->ijab", Y, t1)
    D = eps_o[:, None, None, None] + eps_o[None, :, None, None] - eps_v[None, None, :, None] - eps_v[None, None, None, :]
    small = np.abs(D) < 1e-12
    t22 = np.where(small, 0.0, num / np.where(small, 1.0, D))
    J = G.transpose(0, 2, 1, 3)                       # (ia|jb) as [i,j,a,b]
    K = G.transpose(0, 2, 3, 1)                       # (ib|ja) as [i,j,a,b]
    return float(np.sum((2.0 * J - K) * t22)), t22


def scramble(B):
    """B~[q, r] = L_q(r mod o, r div o) for r = i*v + a: the reference reads
    the column-major flattening of each o x v block L_q as if it were
    row-major (cholesky_ump3.cc:65-97 with the row indices of :228-229,
    :284 and :510)."""
    Q, o, v = B.shape
    r = np.arange(o * v)
    return B.reshape(Q, o * v)[:, (r % o) * v + r // o].reshape(Q, o, v)


def _den(eo_l, ev_l, eo_r, ev_r):
    return (eo_l[:, None, None, None] + eo_r[None, :, None, None] - ev_l[None, None, :, None]
            - ev_r[None, None, None, :])


def _div(x, D):
    ok = np.abs(D) > 1e-12
    return np.where(ok, x / np.where(ok, D, 1.0), 0.0)


def ump3(Ba, Bb, Looa, Loob, Lvva, Lvvb, eoa, eva, eob, evb):
    """(E_ss_aa, E_ss_bb, E_os) of CholeskyUMP3::compute (cholesky_ump3.cc:
    160-637).  Inputs as in rmp3, per spin.  First-order amplitudes from the
    scrambled blocks (:212-301); the three energy functions with their own
    index conventions (:308-637), including the |v| > 1e-12 skips of the
    hole-hole terms."""
    Sa, Sb = scramble(Ba), scramble(Bb)
    Ga = np.einsum("qia,qjb->iajb", Sa, Sa)
    Gb = np.einsum("qia,qjb->iajb", Sb, Sb)
    Gab = np.einsum("qia,qjb->iajb", Sa, Sb)
    taa = _div(Ga.transpose(0, 2, 1, 3) - Ga.transpose(0, 2, 3, 1), _den(eoa, eva, eoa, eva))
    tbb = _div(Gb.transpose(0, 2, 1, 3) - Gb.transpose(0, 2, 3, 1), _den(eob, evb, eob, evb))
    tab = _div(Gab.transpose(0, 2, 1, 3), _den(eoa, eva, eob, evb))

    def same_spin(B, S, Loo, Lvv, t, eo, ev, beta):
        O4 = np.einsum("qkl,qij->klij", Loo, Loo)                      # (kl|ij)
        V = O4 - O4.transpose(0, 1, 3, 2)                               # (kl|ij) - (kl|ji)
        V = np.where(np.abs(V) > 1e-12, V, 0.0)
        new = 0.5 * np.einsum("klij,klab->ijab", V, t)                  # hole-hole
        W = np.einsum("qab,qcd->abcd", Lvv, Lvv)                        # (ab|cd)
        if beta:   # :480-489, 1/2 [(ab|cd) - (ab|dc)] t_cd
            new += 0.5 * (np.einsum("abcd,ijcd->ijab", W, t) - np.einsum("abdc,ijcd->ijab", W, t))
        else:      # :378-411, sum_cd (ab|cd) t_cd
            new += np.einsum("abcd,ijcd->ijab", W, t)
        OV = np.einsum("qkj,qbc->kjbc", Loo, Lvv)                       # (kj|bc)
        G = np.einsum("qkc,qjb->kcjb", S if beta else B, S if beta else B)
        M = OV.transpose(0, 2, 1, 3) - G.transpose(0, 3, 2, 1)          # M[k,b,j,c] = (kj|bc) - (kc|jb)
        new += (np.einsum("kbjc,ikac->ijab", M, t) - np.einsum("kajc,ikbc->ijab", M, t)
                - np.einsum("kbic,jkac->ijab", M, t) + np.einsum("kaic,jkbc->ijab", M, t))
        return 0.25 * float(np.sum(t * _div(new, _den(eo, ev, eo, ev))))

    E_aa = same_spin(Ba, Sa, Looa, Lvva, taa, eoa, eva, False)
    E_bb = same_spin(Bb, Sb, Loob, Lvvb, tbb, eob, evb, True)
    # opposite spin (:537-637)
    O4 = np.einsum("qki,qlj->kilj", Looa, Loob)                         # (ki|lj)
    O4 = np.where(np.abs(O4) > 1e-12, O4, 0.0)
    new = np.einsum("kilj,klab->ijab", O4, tab)
    new += np.einsum("qac,qbd,ijcd->ijab", Lvva, Lvvb, tab, optimize=True)   # (ac|bd) t_cd
    Gaa = np.einsum("qka,qic->kaic", Ba, Ba)                            # (ka|ic)
    Gbb = np.einsum("qlb,qjd->lbjd", Bb, Bb)                            # (lb|jd)
    new += np.einsum("kaic,kjcb->ijab", Gaa, tab) + np.einsum("lbjd,ilad->ijab", Gbb, tab)
    E_os = float(np.sum(tab * _div(new, _den(eoa, eva, eob, evb))))
    return E_aa, E_bb, E_os
