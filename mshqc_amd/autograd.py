"""torch.autograd.Function over the CD-MP2 backward (PyTorch-ROCm only at
this layer; the arithmetic is the HIP library's).

The forward runs the fused energy + gradient pass of the C ABI
(mshqc_hip_rmp2_grad / mshqc_hip_ump2_grad) when any input needs a gradient,
keeps dE/dB and dE/deps, and the backward scales them by the incoming
gradient.  Amplitudes are recomputed per occupied pair block inside that
pass; nothing of size o^2 v^2 is stored.

    E = rmp2_energy(B, eps_o, eps_v, precision=0)   # B: [Q, o, v] float64 cuda
    E.backward()                                     # B.grad = dE/dB = 4 Gamma
"""
from __future__ import annotations

import torch

from . import hip


class _RMP2Energy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, B, eps_o, eps_v, precision):
        c = hip.context(B.device.index or 0)
        need = ctx.needs_input_grad
        if any(need[:3]):
            E, dB, deo, dev = c.rmp2_grad(B.contiguous(), eps_o.contiguous(), eps_v.contiguous(),
                                          precision=precision, want=need[:3])
            ctx.save_for_backward(*(t if t is not None else torch.empty(0, device=B.device)
                                    for t in (dB, deo, dev)))
        else:
            E = c.rmp2_energy(B.contiguous(), eps_o.contiguous(), eps_v.contiguous(), precision=precision)
        return torch.tensor(E, dtype=torch.float64, device=B.device)

    @staticmethod
    def backward(ctx, gE):
        dB, deo, dev = ctx.saved_tensors
        need = ctx.needs_input_grad
        return (gE * dB if need[0] else None, gE * deo if need[1] else None,
                gE * dev if need[2] else None, None)


class _UMP2Energy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Ba, Bb, eoa, eva, eob, evb, precision):
        c = hip.context(Ba.device.index or 0)
        args = [t.contiguous() for t in (Ba, Bb, eoa, eva, eob, evb)]
        if any(ctx.needs_input_grad[:6]):
            E, grads = c.ump2_grad(*args, precision=precision)
            ctx.save_for_backward(*grads)
        else:
            E = c.ump2_energy(*args, precision=precision)
        return torch.tensor(sum(E), dtype=torch.float64, device=Ba.device)

    @staticmethod
    def backward(ctx, gE):
        grads = ctx.saved_tensors
        return tuple(gE * g if need else None for g, need in zip(grads, ctx.needs_input_grad[:6])) + (None,)


def rmp2_energy(B, eps_o, eps_v, precision: int = 0):
    """Differentiable CD-RMP2 correlation energy (0-d float64 tensor)."""
    return _RMP2Energy.apply(B, eps_o, eps_v, precision)


def ump2_energy(Ba, Bb, eps_oa, eps_va, eps_ob, eps_vb, precision: int = 0):
    """Differentiable CD-UMP2 correlation energy E_ss(aa) + E_ss(bb) + E_os."""
    return _UMP2Energy.apply(Ba, Bb, eps_oa, eps_va, eps_ob, eps_vb, precision)
