"""Exact CPU model of the split-BF16 energy (test helper).

Each FP64 element x of B is cut, per Q block of its (i, a) row (128 or 160
wide, as the kernel chooses: 10 stages of 16 Q when that pads Q less than 8
stages do, except at five slices), into
k slices s_p = 2^(e-8) m_p 2^(-9p) with 2^e the block bound and m_p integers,
|m_p| <= 256: m_0 = rint(256 x / 2^e), then m_p = rint(512 r) of each
residual r (|r| <= 1/2), rounding to nearest at each step.  The model
keeps the slice products with p + q <= k - 1 in exact FP64 (the kernel adds
them on FP32 MFMA accumulators, so the two differ only by FP32 rounding of
the p + q >= 2 products) and takes the self terms (ia|ia) of the pairs i = j
as the FP64 sums sum_Q B_ia^2.  The energy is then the CD-RMP2 expression
(cholesky_rmp2.cc:118-155) over all ordered pairs.
"""
import numpy as np

def pair_q_block(Q, k):
    """The kernel's Q block: 16 x 10 if that pads Q to fewer 16-stages than
    16 x 8, except at k = 5 slices, which keep 16 x 8."""
    st = -(-Q // 16)
    return 160 if k != 5 and -(-st // 10) * 10 < -(-st // 8) * 8 else 128


def slices(B, k):
    """k arrays s_p [Q, o, v] whose sum is B up to the truncation."""
    Q = B.shape[0]
    QB = pair_q_block(Q, k)
    Qp = (Q + QB - 1) // QB * QB
    Bp = np.zeros((Qp,) + B.shape[1:])
    Bp[:Q] = B
    Bb = Bp.reshape(Qp // QB, QB, *B.shape[1:])
    mx = np.abs(Bb).max(axis=1, keepdims=True)
    _, e = np.frexp(mx)
    e = np.where(mx > 0, np.maximum(e, -127), 0)
    scale = np.ldexp(1.0, e)
    r = Bb / scale
    out = []
    for p in range(k):
        t = r * (256.0 if p == 0 else 512.0)
        m = np.rint(t)
        out.append((m * np.ldexp(scale, -8 - 9 * p)).reshape(Qp, *B.shape[1:])[:Q])
        r = t - m
    return out


def rmp2_energy(B, eps_o, eps_v, k):
    """The split-BF16 model energy with k slices (pure numpy; small sizes)."""
    Q, o, v = B.shape
    sl = slices(B, k)
    part = [sum(sl[q] for q in range(k - p)).reshape(Q, o * v) for p in range(k)]
    self_terms = np.einsum("qia,qia->ia", B, B)
    E = 0.0
    for i in range(o):
        G = sum(sl[p][:, i, :].T @ part[p] for p in range(k)).reshape(v, o, v)
        G[np.arange(v), i, np.arange(v)] = self_terms[i]
        for j in range(o):
            # the kept set {p + q <= k - 1} is symmetric in p and q, so
            # (ib|ja) of the model is the transpose of its (ia|jb) block
            g = G[:, j, :]
            D = eps_o[i] + eps_o[j] - eps_v[:, None] - eps_v[None, :]
            E += np.sum(g * (2.0 * g - g.T) / D)
    return E
