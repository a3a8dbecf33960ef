"""mshqc_amd: MI355X-native Cholesky-vector hot path of mshqc.

* ``mshqc_amd.hip``        -- ctypes binding of libmshqc_hip.so (C ABI in
                              include/mshqc_hip.h) for torch device tensors
* ``mshqc_amd._mshqc``     -- pybind11 module with the drop-in C++ classes
                              (CholeskyERI, CholeskyRHF, CholeskyUHF,
                              CholeskyRMP2, CholeskyUMP2, ...)
* ``mshqc_amd.autograd``   -- torch.autograd.Function over the MP2 backward
* ``mshqc_amd.synthetic``  -- seeded workload generators
"""
import os

__all__ = ["hip", "synthetic"]

_HERE = os.path.dirname(os.path.abspath(__file__))


def native_module():
    """Import the pybind11 module (raises if it was not built)."""
    from . import _mshqc  # noqa: F401

    return _mshqc
