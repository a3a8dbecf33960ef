"""Writes the full-size fixtures of the plain FP64 workloads (Q=2400) from the
CPU oracle, for the GPU tests in tests/test_gpu_full_size.py.

  cfg1_rmp2_o60_q2400.json   CD-RMP2 energy and dE/deps at Q=2400, o=60,
                             v=540 (mshqc_amd.synthetic workload "mp2_o60")
  cfg1_jk_n600_q2400.json    J and K (general density P = 2 C_occ C_occ^T,
                             with and without the exchange screen) and the
                             AO->MO transform at n=600, Q=2400, o=60
                             (workload "jk600"), as functionals: traces,
                             Frobenius norms, diagonals, the first row and
                             seeded bilinear projections u^T M w

The inputs are regenerated from their seeds (numpy default_rng), so the
fixtures hold outputs only.  Run: python tests/golden/make_cfg1_fixtures.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from mshqc_amd import synthetic as syn  # noqa: E402

NPROJ = 16


def projections(M, seed):
    """u_k^T M w_k for NPROJ seeded unit-normal u, w (M: [n, n])."""
    rng = np.random.default_rng(seed)
    n = M.shape[0]
    U = rng.standard_normal((NPROJ, n))
    W = rng.standard_normal((NPROJ, M.shape[1]))
    return [float(U[k] @ M @ W[k]) for k in range(NPROJ)]


def matrix_functionals(M, seed):
    return {"trace": float(np.trace(M)), "fro": float(np.linalg.norm(M)), "diag": M.diagonal().tolist(),
            "row0": M[0].tolist(), "proj": projections(M, seed), "proj_seed": seed}


def transform_vir(n, o, seed):
    """The virtual block the jk600 workload pairs with C_occ (bench.py)."""
    return np.ascontiguousarray(np.linalg.qr(np.random.default_rng(seed + 1).standard_normal((n, n - o)))[0])


def b_functionals(B, seed):
    """Per-Q projections sum_ia B[q,i,a] u[i,a] (u seeded) and the norm."""
    Q, o, v = B.shape
    u = np.random.default_rng(seed).standard_normal((o, v))
    return {"per_q": np.einsum("qia,ia->q", B, u).tolist(), "fro": float(np.linalg.norm(B)), "proj_seed": seed}


def main():
    t0 = time.time()
    B, eo, ev = syn.workload("mp2_o60")
    E, g, deo, dev = oracle.rmp2_grad(B, eo, ev)
    out = {"workload": "mp2_o60", "Q": B.shape[0], "o": B.shape[1], "v": B.shape[2], "energy": E,
           "dE_deps_o": deo.tolist(), "dE_deps_v": dev.tolist(), "dEdB": b_functionals(g, 501)}
    with open(os.path.join(HERE, "cfg1_rmp2_o60_q2400.json"), "w") as f:
        json.dump(out, f)
    print(f"rmp2 o=60: E = {E!r} ({time.time() - t0:.0f} s)", flush=True)
    del B, g

    w = syn.WORKLOADS["jk600"]
    L, C, P = syn.workload("jk600")
    J, K = oracle.jk(L, P)
    Js, Ks = oracle.jk(L, P, screen=True)
    Cv = transform_vir(w["n"], w["o"], w["seed"])
    Bt = oracle.transform(L, C, Cv)
    out = {"workload": "jk600", "Q": w["Q"], "n": w["n"], "o": w["o"],
           "J": matrix_functionals(J, 601), "K": matrix_functionals(K, 602),
           "J_screen": matrix_functionals(Js, 601), "K_screen": matrix_functionals(Ks, 602),
           "transform": b_functionals(Bt, 603)}
    with open(os.path.join(HERE, "cfg1_jk_n600_q2400.json"), "w") as f:
        json.dump(out, f)
    print(f"jk n=600: tr J {np.trace(J)!r} tr K {np.trace(K)!r} ({time.time() - t0:.0f} s)", flush=True)


if __name__ == "__main__":
    main()
