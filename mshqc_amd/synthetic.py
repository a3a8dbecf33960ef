"""Seeded synthetic inputs for the Cholesky-vector workloads.

The generators reproduce the stated synthetic recipes exactly (numpy
``default_rng`` with a fixed seed per workload):

* ``L^Q_{mu nu}``: ``A = 0.1 * N(0,1) * exp(-|mu-nu|/50)``, ``L = (A + A^T)/2``
* ``C_occ``: Q factor of an ``n x o`` N(0,1) matrix (orthonormal columns)
* ``B^Q_ia = 0.05 * N(0,1) * exp(-2a/v)``, ``a = 0..v-1``
* ``eps_i ~ U[-1.5, -0.3]`` sorted ascending, ``eps_a = 0.1 + 3 (a/v)^1.5``

Named workloads (sizes):

==============  ==========================================================
``jk600``       L: Q=2400, n=600; C_occ 600 x 60      (J/K build)
``mp2_o60``     B: Q=2400, o=60, v=540                 (plain FP64 RMP2)
``mp2_o100``    B: Q=4000, o=100, v=1100               (precision form)
``mp2_o40``     B: Q=1600, o=40, v=360                 (differentiation form)
``ump2_o120``   Q=5600, o_a=120, v_a=1380, o_b=118, v_b=1382 (CD-UMP2)
``omp2_o40``    Q=1600, o_a=40, v_a=360, o_b=38, v_b=362  (CD-OMP2 amplitudes + OPDM)
``cd_n100``     ERI matrix M = L^T L from L [1500, 100*100] (pivoted Cholesky, thr 1e-6)
``mp3_o20``     L: Q=800, n=200; o=20, v=180 (CD-RMP3 from the MO blocks)
``ump3_o20``    L: Q=800, n=200; o_a=20, o_b=18 (CD-UMP3 from the MO blocks)
``transform_o100``  L: Q=4000, n=1200; o=100, v=1100 (AO->MO transform;
                L made on the device by bench.py, not here)
==============  ==========================================================
"""
from __future__ import annotations

import numpy as np

WORKLOADS = {
    "jk600": dict(kind="jk", Q=2400, n=600, o=60, seed=1001),
    "mp2_o60": dict(kind="rmp2", Q=2400, o=60, v=540, seed=1002),
    "mp2_o100": dict(kind="rmp2", Q=4000, o=100, v=1100, seed=1003),
    "mp2_o40": dict(kind="rmp2", Q=1600, o=40, v=360, seed=1004),
    "ump2_o120": dict(kind="ump2", Q=5600, oa=120, va=1380, ob=118, vb=1382, seed=1005),
    "omp2_o40": dict(kind="omp2", Q=1600, oa=40, va=360, ob=38, vb=362, seed=1006),
    "cd_n100": dict(kind="cd", Q=1500, n=100, thr=1e-6, seed=1007),
    "mp3_o20": dict(kind="rmp3", Q=800, n=200, o=20, seed=1008),
    "ump3_o20": dict(kind="ump3", Q=800, n=200, oa=20, ob=18, seed=1009),
    # L generated on the device (bench.py run_transform): 46 GB
    "transform_o100": dict(kind="transform", Q=4000, n=1200, o=100, seed=1010),
}


def gen_L(Q, n, rng, dtype=np.float64):
    """Symmetric L^Q [Q,n,n] with a |mu-nu| decay envelope."""
    idx = np.arange(n)
    env = np.exp(-np.abs(idx[:, None] - idx[None, :]) / 50.0)
    L = np.empty((Q, n, n), dtype=dtype)
    for q in range(Q):
        A = 0.1 * rng.standard_normal((n, n)) * env
        L[q] = 0.5 * (A + A.T)
    return L


def gen_Cocc(n, o, rng):
    """Orthonormal n x o block: Q factor of a Gaussian matrix."""
    Qm, _ = np.linalg.qr(rng.standard_normal((n, o)))
    return np.ascontiguousarray(Qm)


def gen_B(Q, o, v, rng):
    """B [Q,o,v] = 0.05 N(0,1) exp(-2a/v)."""
    env = np.exp(-2.0 * np.arange(v) / v)
    return 0.05 * rng.standard_normal((Q, o, v)) * env[None, None, :]


def gen_eps(o, v, rng):
    """Occupied U[-1.5,-0.3] sorted ascending, virtual 0.1 + 3 (a/v)^1.5."""
    eo = np.sort(rng.uniform(-1.5, -0.3, size=o))
    ev = 0.1 + 3.0 * (np.arange(v) / v) ** 1.5
    return eo, ev


def rmp2_inputs(Q, o, v, seed):
    """(B, eps_o, eps_v) for a closed-shell workload."""
    rng = np.random.default_rng(seed)
    B = gen_B(Q, o, v, rng)
    eo, ev = gen_eps(o, v, rng)
    return B, eo, ev


def ump2_inputs(Q, oa, va, ob, vb, seed):
    """Independent alpha and beta (B, eps_o, eps_v) from one seeded stream."""
    rng = np.random.default_rng(seed)
    Ba = gen_B(Q, oa, va, rng)
    eoa, eva = gen_eps(oa, va, rng)
    Bb = gen_B(Q, ob, vb, rng)
    eob, evb = gen_eps(ob, vb, rng)
    return (Ba, eoa, eva), (Bb, eob, evb)


def jk_inputs(Q, n, o, seed):
    """(L, C_occ, P) with P = 2 C_occ C_occ^T."""
    rng = np.random.default_rng(seed)
    L = gen_L(Q, n, rng)
    C = gen_Cocc(n, o, rng)
    return L, C, 2.0 * C @ C.T


def workload(name):
    """Inputs of a named workload (see module docstring)."""
    w = WORKLOADS[name]
    if w["kind"] == "rmp2":
        return rmp2_inputs(w["Q"], w["o"], w["v"], w["seed"])
    if w["kind"] == "rmp3":
        rng = np.random.default_rng(w["seed"])
        L = gen_L(w["Q"], w["n"], rng)
        C = gen_Cocc(w["n"], w["n"], rng)
        eo, ev = gen_eps(w["o"], w["n"] - w["o"], rng)
        return L, C, eo, ev
    if w["kind"] == "ump3":
        rng = np.random.default_rng(w["seed"])
        n = w["n"]
        L = gen_L(w["Q"], n, rng)
        Ca, Cb = gen_Cocc(n, n, rng), gen_Cocc(n, n, rng)
        eoa, eva = gen_eps(w["oa"], n - w["oa"], rng)
        eob, evb = gen_eps(w["ob"], n - w["ob"], rng)
        return L, (Ca, eoa, eva), (Cb, eob, evb)
    if w["kind"] == "cd":
        rng = np.random.default_rng(w["seed"])
        return gen_L(w["Q"], w["n"], rng).reshape(w["Q"], w["n"] * w["n"])
    if w["kind"] in ("ump2", "omp2"):
        return ump2_inputs(w["Q"], w["oa"], w["va"], w["ob"], w["vb"], w["seed"])
    return jk_inputs(w["Q"], w["n"], w["o"], w["seed"])
