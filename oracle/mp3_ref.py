"""numpy restatement of CD-RMP3 (src/mp3/cholesky_rmp3.cc:79-200) -- TEST
INFRASTRUCTURE ONLY (tiny sizes).

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
