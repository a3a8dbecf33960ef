"""CPU FP64 oracle for the Cholesky-vector hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg may import it; the product (``mshqc_amd``)
never does.  The arithmetic lives in ``cd_oracle.c`` (built by ``make -C
oracle`` into ``liboracle.so``); this module only wraps it with numpy.

Every function names the reference routine it restates (file:line into
syahrulhidayat/mshqc).  Array layouts are C order:
``L[Q,n,n]``, ``B[Q,o,v]``, ``C[n,m]``, ``eri[n,n,n,n]``, ``t2[o,o,v,v]``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def build() -> str:
    """Compile liboracle.so with the committed Makefile (gcc, -O3 -fopenmp)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        lib.orc_rmp2_energy.restype = ctypes.c_double
        lib.orc_rmp2_energy.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D,
                                        _I, ctypes.c_int, _D]
        lib.orc_rmp2_grad.restype = ctypes.c_double
        lib.orc_rmp2_grad.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D,
                                      _I, ctypes.c_int, _D, _D, _D]
        lib.orc_ump2.restype = None
        lib.orc_ump2.argtypes = [_D, _D] + [ctypes.c_int] * 5 + [_D] * 4 + [_D] * 7
        lib.orc_jk.restype = None
        lib.orc_jk.argtypes = [_D, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, ctypes.c_int,
                               _D, _D]
        lib.orc_transform.restype = None
        lib.orc_transform.argtypes = [_D, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, _D,
                                      ctypes.c_int, _D]
        lib.orc_half_transform.restype = None
        lib.orc_half_transform.argtypes = [_D, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, _D]
        lib.orc_cholesky.restype = ctypes.c_int
        lib.orc_cholesky.argtypes = [_D, ctypes.c_int, ctypes.c_double, ctypes.c_int, _D]
        lib.orc_num_threads.restype = ctypes.c_int
        _LIB = lib
    return _LIB


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def num_threads() -> int:
    return _lib().orc_num_threads()


def rmp2_energy(B, eps_o, eps_v, rows=None, want_t2=False):
    """CD-RMP2 correlation energy (cholesky_rmp2.cc:118-155).

    Returns ``E`` or ``(E, t2)``.  ``rows`` restricts the first occupied index
    (timing samples)."""
    B, eo, ev = _c(B), _c(eps_o), _c(eps_v)
    Q, o, v = B.shape
    t2 = np.zeros((o, o, v, v)) if want_t2 else None
    r = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    E = _lib().orc_rmp2_energy(_p(B), Q, o, v, _p(eo), _p(ev),
                               None if r is None else r.ctypes.data_as(_I),
                               0 if r is None else len(r), _p(t2))
    return (E, t2) if want_t2 else E


def rmp2_grad(B, eps_o, eps_v, rows=None):
    """Energy, dE/dB [Q,o,v], dE/deps_occ [o], dE/deps_vir [v] of CD-RMP2."""
    B, eo, ev = _c(B), _c(eps_o), _c(eps_v)
    Q, o, v = B.shape
    g = np.zeros_like(B)
    deo = np.zeros(o)
    dev = np.zeros(v)
    r = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    E = _lib().orc_rmp2_grad(_p(B), Q, o, v, _p(eo), _p(ev),
                             None if r is None else r.ctypes.data_as(_I),
                             0 if r is None else len(r), _p(g), _p(deo), _p(dev))
    return E, g, deo, dev


def ump2(Ba, Bb, eps_oa, eps_va, eps_ob, eps_vb, grad=False):
    """CD-UMP2 (cholesky_ump2.cc:299-465): returns ``(E_ss_aa, E_ss_bb, E_os)``
    and, with ``grad``, ``(dBa, dBb, deoa, deva, deob, devb)``."""
    Ba, Bb = _c(Ba), _c(Bb)
    eoa, eva, eob, evb = map(_c, (eps_oa, eps_va, eps_ob, eps_vb))
    Q, oa, va = Ba.shape
    Qb, ob, vb = Bb.shape
    assert Q == Qb
    E = np.zeros(3)
    if grad:
        dBa, dBb = np.zeros_like(Ba), np.zeros_like(Bb)
        deoa, deva, deob, devb = np.zeros(oa), np.zeros(va), np.zeros(ob), np.zeros(vb)
    else:
        dBa = dBb = deoa = deva = deob = devb = None
    _lib().orc_ump2(_p(Ba), _p(Bb), Q, oa, va, ob, vb, _p(eoa), _p(eva), _p(eob), _p(evb),
                    _p(E), _p(dBa), _p(dBb), _p(deoa), _p(deva), _p(deob), _p(devb))
    if grad:
        return tuple(E), (dBa, dBb, deoa, deva, deob, devb)
    return tuple(E)


def jk(L, P, screen=False):
    """CD J and K (cholesky_rhf.cc:108-136 / cholesky_uhf.cc:81-102).

    ``P`` is ``[n,n]`` or ``[nspin,n,n]``; returns ``J[n,n]`` and ``K`` with the
    shape of ``P``."""
    L = _c(L)
    P = _c(P)
    single = P.ndim == 2
    P3 = P[None] if single else P
    Q, n, _ = L.shape
    J = np.zeros((n, n))
    K = np.zeros_like(P3)
    _lib().orc_jk(_p(L), Q, n, _p(P3), P3.shape[0], int(bool(screen)), _p(J), _p(K))
    return J, (K[0] if single else K)


def transform(L, C_occ, C_vir):
    """B^Q = C_occ^T L^Q C_vir (cholesky_rmp2.cc:80-89), B as [Q,o,v]."""
    L, Co, Cv = _c(L), _c(C_occ), _c(C_vir)
    Q, n, _ = L.shape
    o, v = Co.shape[1], Cv.shape[1]
    B = np.zeros((Q, o, v))
    _lib().orc_transform(_p(L), Q, n, _p(Co), o, _p(Cv), v, _p(B))
    return B


def half_transform(L, C):
    """X^Q = L^Q C as [Q,n,m]."""
    L, C = _c(L), _c(C)
    Q, n, _ = L.shape
    m = C.shape[1]
    X = np.zeros((Q, n, m))
    _lib().orc_half_transform(_p(L), Q, n, _p(C), m, _p(X))
    return X


def cholesky(eri, thr, maxvec=None):
    """Pivoted Cholesky of the ERI matrix (cholesky_eri.cc:40-203).

    Returns ``L[nvec, n*n]``."""
    eri = _c(eri)
    n = eri.shape[0]
    mv = n * n if maxvec is None else maxvec
    L = np.zeros((mv, n * n))
    nv = _lib().orc_cholesky(_p(eri), n, float(thr), mv, _p(L))
    return L[:nv].copy()


def rmp2_energy_pairs_rank1(B, eps_o, eps_v, pairs):
    """The reference's energy loop as written (cholesky_rmp2.cc:118-155:
    rank-1 updates per Cholesky vector, then the epilogue; no t2 store) over
    the listed ordered pairs; returns their summed energy.  For timing."""
    B, eo, ev = _c(B), _c(eps_o), _c(eps_v)
    Q, o, v = B.shape
    pr = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
    fn = _lib().orc_rmp2_energy_pairs_rank1
    fn.restype = ctypes.c_double
    fn.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D, _I, ctypes.c_int]
    return fn(_p(B), Q, o, v, _p(eo), _p(ev), pr.ctypes.data_as(_I), len(pr))


def rmp2_grad_pairs(B, eps_o, eps_v, pairs):
    """Per-pair E+grad work for listed ordered pairs (distinct first indices);
    returns (E over the pairs, partial dE/dB)."""
    B, eo, ev = _c(B), _c(eps_o), _c(eps_v)
    Q, o, v = B.shape
    pr = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
    assert len(set(pr[:, 0].tolist())) == len(pr), "first indices must be distinct"
    g = np.zeros_like(B)
    lib = _lib()
    fn = lib.orc_rmp2_grad_pairs
    fn.restype = ctypes.c_double
    fn.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D, _I, ctypes.c_int, _D]
    E = fn(_p(B), Q, o, v, _p(eo), _p(ev), pr.ctypes.data_as(_I), len(pr), _p(g))
    return E, g
