"""mshqc_hip_rmp2_energy_noalloc: the FP64 CD-RMP2 energy without any device
allocation, against mshqc_hip_rmp2_energy (the same kernels, so bit for bit)
on two calls in a row, the second with a smaller input."""
import ctypes

import numpy as np
import pytest

import oracle
from conftest import to_dev
from mshqc_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def test_noalloc_matches_rmp2_energy(ctx):
    import torch

    E_out = torch.zeros(1, dtype=torch.float64, device="cuda")
    for Q, o, v, seed in [(300, 12, 200, 91), (120, 5, 70, 92)]:
        B, eo, ev = syn.rmp2_inputs(Q, o, v, seed)
        d = to_dev(B, eo, ev)
        E_ref = ctx.rmp2_energy(*d)
        E_out.fill_(float("nan"))
        torch.cuda.synchronize()
        ctx.rmp2_energy_noalloc(*d, E_out)
        torch.cuda.synchronize()
        E = E_out.item()
        print(f"noalloc Q={Q} o={o} v={v}: {E!r} vs {E_ref!r}", flush=True)
        assert E == E_ref
        E_orc = oracle.rmp2_energy(B, eo, ev)
        assert abs(E - E_orc) <= 1e-12 * abs(E_orc)


def test_noalloc_allocates_nothing(ctx):
    """Device free memory is the same before and after a call (after one call
    has loaded the kernels), and the result stays bitwise stable."""
    import torch

    B, eo, ev = syn.rmp2_inputs(200, 8, 150, 93)
    d = to_dev(B, eo, ev)
    E_out = torch.zeros(1, dtype=torch.float64, device="cuda")
    ctx.rmp2_energy_noalloc(*d, E_out)
    torch.cuda.synchronize()
    E1 = E_out.item()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(3):
        ctx.rmp2_energy_noalloc(*d, E_out)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0
    assert E_out.item() == E1


def test_noalloc_rejects_sizes_beyond_its_scratch(ctx):
    """o = 400, v = 2000 needs more than the fixed 64 MiB: a clear error
    before any kernel is queued (the small buffers passed are never read)."""
    import torch

    from mshqc_amd import hip

    lib = hip.lib()
    buf = torch.zeros(16, dtype=torch.float64, device="cuda")
    p = ctypes.c_void_p(buf.data_ptr())
    rc = lib.mshqc_hip_rmp2_energy_noalloc(ctx.h, p, 10, 400, 2000, p, p, p)
    msg = lib.mshqc_hip_last_error(ctx.h).decode()
    print(msg, flush=True)
    assert rc == 1 and "scratch" in msg
