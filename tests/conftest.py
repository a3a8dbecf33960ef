import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running check")


@pytest.fixture(scope="session")
def ctx():
    """The HIP context used by GPU tests (fails loudly if the library is missing)."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from mshqc_amd import hip

    return hip.context(0)


def to_dev(*arrays):
    import torch

    return [torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64).cuda() for a in arrays]


def to_np(t):
    return t.detach().cpu().numpy()


def check_close(got, ref, rtol, name="value", atol=0.0):
    """Max-norm relative check with a short failure message (no array dumps)."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    if got.shape != ref.shape:
        raise AssertionError(f"{name}: shape {got.shape} != {ref.shape}")
    scale = float(np.abs(ref).max()) if ref.size else 0.0
    err = float(np.abs(got - ref).max()) if ref.size else 0.0
    if not (err <= rtol * scale + atol):
        idx = np.unravel_index(np.argmax(np.abs(got - ref)), ref.shape) if ref.ndim else ()
        raise AssertionError(f"{name}: max |diff| {err:.3e} > {rtol:g} * {scale:.3e} + {atol:g} "
                             f"(worst at {idx})")
    return err
