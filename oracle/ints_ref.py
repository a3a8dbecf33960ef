"""Independent primitive-integral checker -- TEST INFRASTRUCTURE ONLY.

s-type Gaussian integrals in closed form (overlap, kinetic, nuclear
attraction, ERI with the Boys function F0(T) = sqrt(pi/T) erf(sqrt T) / 2),
and p-type integrals from them by differentiation with respect to a centre:
    (x - A_x) exp(-a |r - A|^2) = (1 / 2a) d/dA_x exp(-a |r - A|^2),
taken by central differences.  This shares no code or recursion with the
McMurchie-Davidson engine it checks.
"""
from __future__ import annotations

import math

import numpy as np


def _F0(T):
    return 1.0 if T < 1e-14 else 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))


def s_overlap(a, A, b, B):
    p = a + b
    R2 = float(np.sum((np.asarray(A) - np.asarray(B)) ** 2))
    return (math.pi / p) ** 1.5 * math.exp(-a * b / p * R2)


def s_kinetic(a, A, b, B):
    mu = a * b / (a + b)
    R2 = float(np.sum((np.asarray(A) - np.asarray(B)) ** 2))
    return mu * (3.0 - 2.0 * mu * R2) * s_overlap(a, A, b, B)


def s_nuclear(a, A, b, B, C, Z):
    p = a + b
    A, B, C = map(np.asarray, (A, B, C))
    P = (a * A + b * B) / p
    R2 = float(np.sum((A - B) ** 2))
    return -Z * 2 * math.pi / p * math.exp(-a * b / p * R2) * _F0(p * float(np.sum((P - C) ** 2)))


def s_eri(a, A, b, B, c, C, d, D):
    A, B, C, D = map(np.asarray, (A, B, C, D))
    p, q = a + b, c + d
    P, Qc = (a * A + b * B) / p, (c * C + d * D) / q
    K = math.exp(-a * b / p * float(np.sum((A - B) ** 2)) - c * d / q * float(np.sum((C - D) ** 2)))
    al = p * q / (p + q)
    return 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * K * _F0(al * float(np.sum((P - Qc) ** 2)))


def with_p(fn, exps, centers, comps, h=1e-4):
    """Integral fn(*exps/centers) with Cartesian p components comps[i] in
    (None, 0, 1, 2) on the centres, by central differences (up to two p)."""
    idx = [i for i, c in enumerate(comps) if c is not None]
    if not idx:
        return fn(exps, centers)

    def shifted(sh):
        cs = [np.array(c, dtype=float) for c in centers]
        for i, s in sh:
            cs[i][comps[i]] += s
        return fn(exps, cs)

    if len(idx) == 1:
        i = idx[0]
        d = (shifted([(i, h)]) - shifted([(i, -h)])) / (2 * h)
        return d / (2 * exps[i])
    i, j = idx
    d = (shifted([(i, h), (j, h)]) - shifted([(i, h), (j, -h)]) - shifted([(i, -h), (j, h)])
         + shifted([(i, -h), (j, -h)])) / (4 * h * h)
    return d / (4 * exps[i] * exps[j])
