"""The plain FP64 and UMP2 workloads at their full stated sizes.

  * Q=2400, o=60, v=540 CD-RMP2 (workload "mp2_o60"): the FP64 energy against
    the CPU oracle's (fixture tests/golden/cfg1_rmp2_o60_q2400.json, written by
    tests/golden/make_cfg1_fixtures.py from the same seeded inputs) to 2 ulp.
    The plain form's bar is 1e-10 Eh, but |E| = 8.0e5 Eh here, where one ulp
    is 1.16e-10 Eh: two summation orders of the same FP64 terms cannot agree
    to better than a few ulp, so the test states the bar in ulp.
  * n=600, Q=2400 J/K (general density P = 2 C_occ C_occ^T with and without
    the exchange screen, and the occupied-factor form) and the AO->MO
    transform to o=60, v=540 (workload "jk600"): against the oracle's
    functionals (traces, norms, diagonals, first rows, seeded projections) in
    tests/golden/cfg1_jk_n600_q2400.json at 1e-12 relative.
  * oa=120/va=1380/ob=118/vb=1382/Q=5600 CD-UMP2 E+grad (workload
    "ump2_o120"): the FP64 path's size-independent identities
    <B, dE/dB> = 4E, sum dE/deps = 0, sum eps dE/deps = -E to 1e-12, and the
    four-slice split-BF16 form within 1 uEh per energy component and 1e-6
    relative per gradient of the FP64 path.
Run time on one MI355X: about 2.5 minutes, most of it numpy generating the
inputs on the host (2.7e9 normal deviates).
"""
import json
import os

import numpy as np
import pytest

from conftest import ROOT, to_np
from mshqc_amd import synthetic as syn

pytestmark = pytest.mark.gpu
GOLD = os.path.join(ROOT, "tests", "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


def _b_functionals(B, seed):
    """Per-Q projections with the fixture's seeded weights, and the norm."""
    Q, o, v = B.shape
    u = np.random.default_rng(seed).standard_normal((o, v))
    return np.einsum("qia,ia->q", B, u), float(np.linalg.norm(B))


def _matrix_check(M, ref, rtol, name):
    rng = np.random.default_rng(ref["proj_seed"])
    n = M.shape[0]
    U = rng.standard_normal((len(ref["proj"]), n))
    W = rng.standard_normal((len(ref["proj"]), M.shape[1]))
    proj = [float(U[k] @ M @ W[k]) for k in range(len(ref["proj"]))]
    scale = float(np.linalg.norm(M))
    assert abs(np.trace(M) - ref["trace"]) <= rtol * scale, name
    assert abs(scale - ref["fro"]) <= rtol * scale, name
    assert _rel(M.diagonal(), ref["diag"]) <= rtol, name
    assert _rel(M[0], ref["row0"]) <= rtol, name
    # a projection u^T M w is a sum of n^2 terms of size |M| |u| |w| / n^2
    pscale = scale * np.linalg.norm(U, axis=1) * np.linalg.norm(W, axis=1)
    assert np.all(np.abs(np.array(proj) - np.array(ref["proj"])) <= rtol * pscale), name


@pytest.mark.timeout(600)
def test_plain_rmp2_q2400_o60_v540(ctx):
    import torch

    ref = _load("cfg1_rmp2_o60_q2400.json")
    B, eo, ev = syn.workload("mp2_o60")
    assert B.shape == (ref["Q"], ref["o"], ref["v"])
    d = [torch.from_numpy(x).cuda() for x in (B, eo, ev)]
    E = ctx.rmp2_energy(*d)
    ulp = np.spacing(abs(ref["energy"]))
    print(f"Q=2400 o=60 v=540: E {E!r} oracle {ref['energy']!r} diff {(E - ref['energy']) / ulp:.1f} ulp "
          f"(ulp {ulp:.2e} Eh)", flush=True)
    assert abs(E - ref["energy"]) <= 2 * ulp
    E2, g, deo, dev = ctx.rmp2_grad(*d)
    assert abs(E2 - ref["energy"]) <= 2 * ulp
    assert _rel(to_np(deo), ref["dE_deps_o"]) <= 1e-12
    assert _rel(to_np(dev), ref["dE_deps_v"]) <= 1e-12
    per_q, fro = _b_functionals(to_np(g), ref["dEdB"]["proj_seed"])
    assert abs(fro - ref["dEdB"]["fro"]) <= 1e-12 * fro
    assert _rel(per_q, ref["dEdB"]["per_q"]) <= 1e-12
    # the split-BF16 form (four slices) within the precision bar
    E4, g4, _, _ = ctx.rmp2_grad(*d, precision=4)
    print(f"  four slices: dE {E4 - E:+.2e} Eh, dE/dB rel {(torch.linalg.norm(g4 - g) / torch.linalg.norm(g)).item():.1e}",
          flush=True)
    assert abs(E4 - E) <= 1e-6
    assert (torch.linalg.norm(g4 - g) / torch.linalg.norm(g)).item() <= 1e-6


@pytest.mark.timeout(600)
def test_jk_and_transform_n600_q2400(ctx):
    import torch

    ref = _load("cfg1_jk_n600_q2400.json")
    w = syn.WORKLOADS["jk600"]
    L, C, P = syn.workload("jk600")
    assert (L.shape[0], L.shape[1], C.shape[1]) == (ref["Q"], ref["n"], ref["o"])
    dL = torch.from_numpy(L).cuda()
    del L
    dC, dP = torch.from_numpy(C).cuda(), torch.from_numpy(P).cuda()
    print(f"J/K n={ref['n']} Q={ref['Q']} o={ref['o']}", flush=True)
    for screen, jn, kn in ((False, "J", "K"), (True, "J_screen", "K_screen")):
        J, K = ctx.jk(dL, dP, screen=screen)
        _matrix_check(to_np(J), ref[jn], 1e-12, jn)
        _matrix_check(to_np(K), ref[kn], 1e-12, kn)
    # occupied-factor form: Y = sqrt(2) C_occ gives the same density P
    Jf, Kf = ctx.jk_factor(dL, (2.0 ** 0.5) * dC)
    _matrix_check(to_np(Jf), ref["J"], 1e-12, "J (factor)")
    _matrix_check(to_np(Kf), ref["K"], 1e-12, "K (factor)")
    n, o = w["n"], w["o"]
    Cv = np.ascontiguousarray(np.linalg.qr(np.random.default_rng(w["seed"] + 1).standard_normal((n, n - o)))[0])
    B = to_np(ctx.transform(dL, dC, torch.from_numpy(Cv).cuda()))
    per_q, fro = _b_functionals(B, ref["transform"]["proj_seed"])
    assert abs(fro - ref["transform"]["fro"]) <= 1e-12 * fro
    assert _rel(per_q, ref["transform"]["per_q"]) <= 1e-12


@pytest.mark.timeout(900)
def test_ump2_full_size(ctx):
    import torch

    (Ba, eoa, eva), (Bb, eob, evb) = syn.workload("ump2_o120")
    d = [torch.from_numpy(x).cuda() for x in (Ba, Bb, eoa, eva, eob, evb)]
    del Ba, Bb
    Ba, Bb, eoa, eva, eob, evb = d
    print(f"ump2 Q={Ba.shape[0]} oa={Ba.shape[1]} va={Ba.shape[2]} ob={Bb.shape[1]} vb={Bb.shape[2]}",
          flush=True)
    E, (dBa, dBb, doa, dva, dob, dvb) = ctx.ump2_grad(*d)
    Et = sum(E)
    assert abs((Ba * dBa).sum().item() + (Bb * dBb).sum().item() - 4 * Et) <= 1e-12 * abs(Et)
    ds = [doa, dva, dob, dvb]
    assert abs(sum(x.sum().item() for x in ds)) <= 1e-12 * abs(Et)
    assert abs(sum((e * x).sum().item() for e, x in zip([eoa, eva, eob, evb], ds)) + Et) <= 1e-12 * abs(Et)
    E4, (gBa, gBb, *gd) = ctx.ump2_grad(*d, precision=4)
    print(f"  four slices: dE {[f'{a - b:+.2e}' for a, b in zip(E4, E)]} Eh", flush=True)
    for got, want in zip(E4, E):
        assert abs(got - want) <= 1e-6
    for got, want in ((gBa, dBa), (gBb, dBb), (torch.cat(gd), torch.cat(ds))):
        rel = (torch.linalg.norm(got - want) / torch.linalg.norm(want)).item()
        print(f"  gradient rel {rel:.2e}", flush=True)
        assert rel <= 1e-6
