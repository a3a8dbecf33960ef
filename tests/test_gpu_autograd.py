"""The backward through torch.autograd at the differentiation workload
(Q=1600, o=40, v=360): central finite differences (step 1e-5) on 64 random
entries of B and on orbital energies, the homogeneity identities, and the
repository's own OMP2 one-particle-density diagonals."""
import numpy as np
import pytest

from conftest import to_dev, to_np
from mshqc_amd import synthetic as syn

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def o40_inputs():
    import torch

    B, eo, ev = syn.workload("mp2_o40")
    return [torch.from_numpy(x).cuda() for x in (B, eo, ev)]


# Central differences of the FP64 energy at step h = 1e-5.  An energy
# evaluation is a fixed-order sum of ~1e8 FP64 terms, and a perturbed input
# changes the rounding of many of them: the test takes that noise as at most
# 8 ulp of E per evaluation (about 1 ulp was measured), so a difference
# quotient resolves a gradient entry to NOISE = 8 ulp(E) / h (1.2e-5 at
# |E| = 1.07e5 Eh here).  The
# truncation error of the central difference, h^2 E'''/6, is below 1e-9 of
# these entries.  Every entry must agree to 1e-6 relative plus that floor.
FD_H = 1e-5


def _noise(E):
    return 8 * np.spacing(abs(E)) / FD_H


@pytest.mark.parametrize("precision", [0, 4])
def test_fd_64_entries(ctx, o40_inputs, precision):
    from mshqc_amd.autograd import rmp2_energy

    B, eo, ev = [t.clone().requires_grad_(True) for t in o40_inputs]
    E = rmp2_energy(B, eo, ev, precision)
    E.backward()
    g = B.grad
    rng = np.random.default_rng(64)
    Q, o, v = B.shape
    h = FD_H
    Bd = o40_inputs[0].clone()
    fds, gs = [], []
    for _ in range(64):
        idx = (int(rng.integers(Q)), int(rng.integers(o)), int(rng.integers(v)))
        x0 = Bd[idx].item()
        Bd[idx] = x0 + h
        Ep = ctx.rmp2_energy(Bd, o40_inputs[1], o40_inputs[2])
        Bd[idx] = x0 - h
        Em = ctx.rmp2_energy(Bd, o40_inputs[1], o40_inputs[2])
        Bd[idx] = x0
        fds.append((Ep - Em) / (2 * h))
        gs.append(g[idx].item())
    fds, gs = np.array(fds), np.array(gs)
    noise = _noise(E.item())
    excess = np.abs(fds - gs) / (1e-6 * np.abs(gs) + noise)
    k = int(np.argmax(excess))
    print(f"precision {precision}: E {E.item()!r}, worst entry |fd - g| = {abs(fds[k] - gs[k]):.2e} "
          f"(g {gs[k]:.3e}, floor {noise:.1e}), {int(np.sum(np.abs(gs) < noise / 1e-6))} of 64 entries "
          f"below 1e6 x the floor", flush=True)
    assert np.all(excess <= 1.0)


def test_fd_every_orbital_energy(ctx, o40_inputs):
    """dE/deps for all 40 occupied and 360 virtual orbital energies against
    central differences, each to 1e-6 relative plus the noise floor; the
    four-slice form's dE/deps within 1e-6 relative of the FP64 one."""
    B, eo, ev = o40_inputs
    E, _, d_o, d_v = ctx.rmp2_grad(B, eo, ev)
    noise = _noise(E)
    for eps, d, other, left in ((eo, d_o, ev, True), (ev, d_v, eo, False)):
        fd = np.empty(eps.shape[0])
        e = eps.clone()
        for k in range(eps.shape[0]):
            x0 = e[k].item()
            e[k] = x0 + FD_H
            Ep = ctx.rmp2_energy(B, e, other) if left else ctx.rmp2_energy(B, other, e)
            e[k] = x0 - FD_H
            Em = ctx.rmp2_energy(B, e, other) if left else ctx.rmp2_energy(B, other, e)
            e[k] = x0
            fd[k] = (Ep - Em) / (2 * FD_H)
        d = to_np(d)
        excess = np.abs(fd - d) / (1e-6 * np.abs(d) + noise)
        print(f"{'occupied' if left else 'virtual'} eps: worst {excess.max():.2f} of the bound", flush=True)
        assert np.all(excess <= 1.0)
    _, _, d_o4, d_v4 = ctx.rmp2_grad(B, eo, ev, precision=4)
    for a, b in ((d_o4, d_o), (d_v4, d_v)):
        assert np.abs(to_np(a) - to_np(b)).max() <= 1e-6 * np.abs(to_np(b)).max()


def test_identities_at_full_size(ctx, o40_inputs):
    B, eo, ev = o40_inputs
    E, g, d_o, d_v = ctx.rmp2_grad(B, eo, ev)
    assert abs((B * g).sum().item() - 4 * E) <= 1e-12 * abs(E)
    assert abs(d_o.sum().item() + d_v.sum().item()) <= 1e-12 * abs(E)
    assert abs((eo * d_o).sum().item() + (ev * d_v).sum().item() + E) <= 1e-12 * abs(E)
    E2 = ctx.rmp2_energy(B, eo, ev)
    assert abs(E2 - E) <= 1e-13 * abs(E)


def test_q_rotation_covariance(ctx):
    import torch

    B, eo, ev = syn.rmp2_inputs(96, 5, 100, 91)
    R, _ = np.linalg.qr(np.random.default_rng(92).standard_normal((96, 96)))
    Br = np.einsum("pq,qia->pia", R, B)
    E1, g1, do1, dv1 = ctx.rmp2_grad(*to_dev(B, eo, ev))
    E2, g2, do2, dv2 = ctx.rmp2_grad(*to_dev(Br, eo, ev))
    assert abs(E1 - E2) <= 1e-12 * abs(E1)
    g1r = np.einsum("pq,qia->pia", R, to_np(g1))
    assert np.abs(g1r - to_np(g2)).max() <= 1e-11 * np.abs(g1r).max()


def test_deps_equals_omp2_opdm_diagonals(ctx):
    from oracle import omp2_ref

    (Ba, eoa, eva), (Bb, eob, evb) = syn.ump2_inputs(40, 3, 11, 2, 12, 93)
    _, grads = ctx.ump2_grad(*to_dev(Ba, Bb, eoa, eva, eob, evb))
    taa, tbb, tab = omp2_ref.t2_blocks(Ba, Bb, eoa, eva, eob, evb)
    Goo_a, Gvv_a, Goo_b, Gvv_b = omp2_ref.opdm(taa, tbb, tab)
    for got, G in zip(grads[2:], (Goo_a, Gvv_a, Goo_b, Gvv_b)):
        assert np.abs(to_np(got) - np.diag(G)).max() <= 1e-12 * np.abs(np.diag(G)).max()


def test_ump2_autograd(ctx):
    import torch

    from mshqc_amd.autograd import ump2_energy

    (Ba, eoa, eva), (Bb, eob, evb) = syn.ump2_inputs(50, 3, 20, 2, 21, 94)
    ts = [t.requires_grad_(True) for t in to_dev(Ba, Bb, eoa, eva, eob, evb)]
    E = ump2_energy(*ts)
    E.backward()
    h = 1e-5
    Bd = ts[0].detach().clone()
    x0 = Bd[3, 1, 5].item()
    Bd[3, 1, 5] = x0 + h
    Ep = sum(ctx.ump2_energy(Bd, *[t.detach() for t in ts[1:]]))
    Bd[3, 1, 5] = x0 - h
    Em = sum(ctx.ump2_energy(Bd, *[t.detach() for t in ts[1:]]))
    fd = (Ep - Em) / (2 * h)
    assert abs(fd - ts[0].grad[3, 1, 5].item()) <= 1e-6 * max(abs(fd), 1e-3)
