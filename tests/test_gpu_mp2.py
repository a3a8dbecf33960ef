"""GPU parity of the MP2 kernels (through the C ABI) against the FP64 oracle.

Bars: the FP64 form agrees with the oracle to 1e-12 relative on energies and
gradients (the rounding of two different summation orders); at larger sizes
size-independent identities are checked instead of the full oracle."""
import numpy as np
import pytest

import oracle
from conftest import check_close, to_dev, to_np
from mshqc_amd import synthetic as syn

pytestmark = pytest.mark.gpu

# (Q, o, v, seed): odd and tile-ragged virtual counts, o = 1, tiny Q.
# (Q, o, v, seed): odd v runs the register-staged FP64 pair kernel, even v the
# direct-to-LDS one (stages of 12 Q; Q = 7 is a single partial stage)
CASES = [(50, 5, 45, 11), (7, 1, 3, 12), (130, 7, 130, 13), (64, 3, 64, 14), (33, 4, 129, 15), (7, 2, 66, 16)]


@pytest.mark.parametrize("Q,o,v,seed", CASES)
def test_rmp2_energy_matches_oracle(ctx, Q, o, v, seed):
    B, eo, ev = syn.rmp2_inputs(Q, o, v, seed)
    E_ref = oracle.rmp2_energy(B, eo, ev)
    dB, deo, dev = to_dev(B, eo, ev)
    E = ctx.rmp2_energy(dB, deo, dev)
    print(f"rmp2 energy Q={Q} o={o} v={v}: gpu {E!r} oracle {E_ref!r}", flush=True)
    assert abs(E - E_ref) <= 1e-12 * abs(E_ref) + 1e-15


@pytest.mark.parametrize("Q,o,v,seed", CASES)
def test_rmp2_grad_matches_oracle(ctx, Q, o, v, seed):
    B, eo, ev = syn.rmp2_inputs(Q, o, v, seed)
    E_ref, g_ref, deo_ref, dev_ref = oracle.rmp2_grad(B, eo, ev)
    dB, deo, dev = to_dev(B, eo, ev)
    E, g, d_o, d_v = ctx.rmp2_grad(dB, deo, dev)
    g, d_o, d_v = to_np(g), to_np(d_o), to_np(d_v)
    print(f"rmp2 grad Q={Q} o={o} v={v}: E {E!r} vs {E_ref!r}", flush=True)
    assert abs(E - E_ref) <= 1e-12 * abs(E_ref) + 1e-15
    check_close(g, g_ref, 1e-12, "g_ref")
    check_close(d_o, deo_ref, 1e-12, "deo_ref")
    check_close(d_v, dev_ref, 1e-12, "dev_ref")


def test_rmp2_grad_small_zblock_matches(ctx):
    """Several occupied block pairs (off-diagonal and diagonal) per call."""
    from mshqc_amd import hip

    B, eo, ev = syn.rmp2_inputs(40, 9, 70, 21)
    E_ref, g_ref, deo_ref, dev_ref = oracle.rmp2_grad(B, eo, ev)
    c = hip.Context(0)
    try:
        c.set_zblock(2 * 70 * 2 * 70)  # two occupied orbitals per block
        dB, deo, dev = to_dev(B, eo, ev)
        E, g, d_o, d_v = c.rmp2_grad(dB, deo, dev)
    finally:
        c.close()
    assert abs(E - E_ref) <= 1e-12 * abs(E_ref)
    check_close(to_np(g), g_ref, 1e-12, "g_ref")
    check_close(to_np(d_o), deo_ref, 1e-12, "deo_ref")
    check_close(to_np(d_v), dev_ref, 1e-12, "dev_ref")


def test_rmp2_t2_matches_oracle(ctx):
    B, eo, ev = syn.rmp2_inputs(30, 4, 70, 31)
    _, t2_ref = oracle.rmp2_energy(B, eo, ev, want_t2=True)
    t2 = to_np(ctx.rmp2_t2(*to_dev(B, eo, ev)))
    check_close(t2, t2_ref, 1e-12, "t2_ref")


def test_rmp2_skip_branch(ctx):
    """|D| < 1e-12 terms are skipped exactly as the reference does."""
    B, eo, ev = syn.rmp2_inputs(20, 3, 10, 41)
    ev = ev.copy()
    eo = eo.copy()
    eo[1] = 0.2
    ev[2] = 0.2  # D_{11,22} = 0
    E_ref, g_ref, deo_ref, dev_ref = oracle.rmp2_grad(B, eo, ev)
    E, g, d_o, d_v = ctx.rmp2_grad(*to_dev(B, eo, ev))
    assert np.isfinite(E) and abs(E - E_ref) <= 1e-12 * abs(E_ref)
    check_close(to_np(g), g_ref, 1e-12, "g_ref")


UCASES = [(40, 3, 20, 4, 17, 51), (70, 5, 66, 4, 70, 52), (16, 1, 5, 2, 3, 53), (9, 3, 40, 2, 42, 54)]


@pytest.mark.parametrize("Q,oa,va,ob,vb,seed", UCASES)
def test_ump2_matches_oracle(ctx, Q, oa, va, ob, vb, seed):
    (Ba, eoa, eva), (Bb, eob, evb) = syn.ump2_inputs(Q, oa, va, ob, vb, seed)
    E_ref, grads_ref = oracle.ump2(Ba, Bb, eoa, eva, eob, evb, grad=True)
    dev = to_dev(Ba, Bb, eoa, eva, eob, evb)
    E = ctx.ump2_energy(*dev)
    print(f"ump2 Q={Q} oa={oa} va={va} ob={ob} vb={vb}: {E} vs {E_ref}", flush=True)
    for x, y in zip(E, E_ref):
        assert abs(x - y) <= 1e-12 * abs(y) + 1e-15
    E2, grads = ctx.ump2_grad(*dev)
    for x, y in zip(E2, E_ref):
        assert abs(x - y) <= 1e-12 * abs(y) + 1e-15
    for a, b in zip(grads, grads_ref):
        a = to_np(a)
        check_close(a, b, 1e-12, "b")


@pytest.mark.parametrize("ob,vb", [(0, 30), (3, 0), (1, 30)])
@pytest.mark.parametrize("precision", [0, 4])
def test_ump2_empty_or_single_beta(ctx, ob, vb, precision):
    """A beta spin with no occupied (ob = 0) or no virtual (vb = 0) orbitals,
    or a single one (no same-spin pair): its blocks give 0 and the alpha
    same-spin energy and gradients are unchanged, as in the reference's loops
    over empty ranges (cholesky_ump2.cc:299-465)."""
    import torch

    (Ba, eoa, eva), (Bb, eob, evb) = syn.ump2_inputs(50, 4, 30, max(ob, 1), max(vb, 1), 54)
    Bb, eob, evb = Bb[:, :ob, :vb], eob[:ob], evb[:vb]
    E_ref, grads_ref = oracle.ump2(Ba, Bb, eoa, eva, eob, evb, grad=True) if ob and vb else (None, None)
    dev = to_dev(Ba, Bb, eoa, eva, eob, evb)
    E = ctx.ump2_energy(*dev, precision=precision)
    E2, grads = ctx.ump2_grad(*dev, precision=precision)
    Eaa = oracle.ump2(Ba, Ba[:, :0, :], eoa, eva, eoa[:0], eva)[0] if E_ref is None else E_ref[0]
    tol = 1e-12 if precision == 0 else 1e-10
    for got in (E, E2):
        assert abs(got[0] - Eaa) <= tol * abs(Eaa)
        assert got[1] == 0.0
        if E_ref is None:
            assert got[2] == 0.0
        else:
            assert abs(got[2] - E_ref[2]) <= tol * abs(E_ref[2])
    assert tuple(grads[1].shape) == (50, ob, vb)
    if ob == 0 or vb == 0:
        _, g_aa = oracle.ump2(Ba, Ba[:, :0, :], eoa, eva, eoa[:0], eva, grad=True)
        check_close(to_np(grads[0]), g_aa[0], 1e-6 if precision else 1e-12, "dE/dBa")


@pytest.mark.parametrize("Q,oa,va,ob,vb,seed", UCASES + [(30, 3, 70, 0, 72, 55)])
def test_ump2_t2_matches_oracle(ctx, Q, oa, va, ob, vb, seed):
    """mshqc_hip_ump2_t2 against g/D from oracle-transformed inputs (the
    reference's compute_t2_amplitudes, cholesky_ump2.cc:471-556: every ordered
    pair, not antisymmetrised), and the blocks' own energies: E_os = sum t_ab g,
    and the same-spin 1/4 sum (g - g^T)(t - t^T) over all i, j."""
    (Ba, eoa, eva), (Bb, eob, evb) = syn.ump2_inputs(Q, oa, va, max(ob, 1), vb, seed)
    Bb, eob = Bb[:, :ob], eob[:ob]
    taa, tbb, tab = (to_np(t) for t in ctx.ump2_t2(*to_dev(Ba, Bb, eoa, eva, eob, evb)))
    E_ref = oracle.ump2(Ba, Bb, eoa, eva, eob, evb)
    for t, X, Y, ex, ey, vx, vy, name in ((taa, Ba, Ba, eoa, eoa, eva, eva, "aa"), (tbb, Bb, Bb, eob, eob, evb, evb, "bb"),
                                          (tab, Ba, Bb, eoa, eob, eva, evb, "ab")):
        if t.size == 0:
            continue
        g = np.einsum("qia,qjb->ijab", X, Y)
        D = ex[:, None, None, None] + ey[None, :, None, None] - vx[None, None, :, None] - vy[None, None, None, :]
        check_close(t, g / D, 1e-12, f"t2_{name}")
        if name == "ab":
            assert abs(np.sum(t * g) - E_ref[2]) <= 1e-12 * abs(E_ref[2])
        else:
            Ess = 0.25 * np.sum((g - g.transpose(0, 1, 3, 2)) * (t - t.transpose(0, 1, 3, 2)))
            assert abs(Ess - E_ref[0 if name == "aa" else 1]) <= 1e-12 * abs(Ess) + 1e-15


def test_ump2_identical_spins_equals_rmp2(ctx):
    B, eo, ev = syn.rmp2_inputs(60, 6, 50, 61)
    dB, deo, dev = to_dev(B, eo, ev)
    Eu = ctx.ump2_energy(dB, dB, deo, dev, deo, dev)
    Er = ctx.rmp2_energy(dB, deo, dev)
    assert abs(sum(Eu) - Er) <= 1e-12 * abs(Er)


def test_rmp2_full_q_v_reduced_o(ctx):
    """Full Q and v of the o=60 workload (Q=2400, v=540) with o=6."""
    B, eo, ev = syn.rmp2_inputs(2400, 6, 540, 1002)
    E_ref = oracle.rmp2_energy(B, eo, ev)
    E = ctx.rmp2_energy(*to_dev(B, eo, ev))
    print(f"rmp2 Q=2400 o=6 v=540: {E!r} vs {E_ref!r} diff {E - E_ref:.3e}", flush=True)
    assert abs(E - E_ref) <= 1e-10
