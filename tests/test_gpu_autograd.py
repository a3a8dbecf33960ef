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


@pytest.mark.parametrize("precision", [0, 5])
def test_fd_64_entries(ctx, o40_inputs, precision):
    import torch

    from mshqc_amd.autograd import rmp2_energy

    B, eo, ev = [t.clone().requires_grad_(True) for t in o40_inputs]
    E = rmp2_energy(B, eo, ev, precision)
    E.backward()
    g = B.grad
    rng = np.random.default_rng(64)
    Q, o, v = B.shape
    h = 1e-5
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
    # Norm-wise over the sample: the difference quotient of an energy of size
    # ~1e5 carries ~1e-11 Eh of rounding noise per evaluation, so entries with
    # tiny gradients are not individually resolvable at step 1e-5.
    rel = np.abs(fds - gs).max() / np.abs(gs).max()
    k = int(np.argmax(np.abs(fds - gs)))
    print(f"precision {precision}: E {E.item()!r}, FD max|diff|/max|g| {rel:.2e} "
          f"(worst fd {fds[k]!r} g {gs[k]!r})", flush=True)
    assert rel < 1e-6
    # orbital energies
    for k in (0, 17, 39):
        e0 = o40_inputs[1].clone()
        e0[k] += h
        Ep = ctx.rmp2_energy(o40_inputs[0], e0, o40_inputs[2])
        e0[k] -= 2 * h
        Em = ctx.rmp2_energy(o40_inputs[0], e0, o40_inputs[2])
        fd = (Ep - Em) / (2 * h)
        assert abs(fd - eo.grad[k].item()) <= 1e-5 * abs(fd)


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
