"""C-ABI surface checks that need no GPU: the library loads and exports every
symbol include/mshqc_hip.h declares."""
import os
import re

import pytest

from conftest import ROOT


def _declared_symbols():
    with open(os.path.join(ROOT, "include", "mshqc_hip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(mshqc_hip_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_table():
    from mshqc_amd import hip

    assert _declared_symbols() == hip.EXPORTED


def test_library_exports_every_declared_symbol():
    from mshqc_amd import hip

    path = hip.LIB_PATH
    if not os.path.exists(path):
        pytest.fail("libmshqc_hip.so not built (run __graft_entry__.build())")
    lib = hip.load_library(path)
    for name in _declared_symbols():
        assert hasattr(lib, name), name
    assert lib.mshqc_hip_version().decode().startswith("mshqc_hip gfx950")


def test_library_has_no_oracle_dependency():
    """The product library must not link or name the CPU oracle."""
    from mshqc_amd import hip

    with open(hip.LIB_PATH, "rb") as f:
        blob = f.read()
    assert b"orc_" not in blob and b"liboracle" not in blob


def test_null_context_is_rejected():
    import ctypes

    from mshqc_amd import hip

    lib = hip.load_library()
    E = ctypes.c_double()
    rc = lib.mshqc_hip_rmp2_energy(None, None, 1, 1, 1, None, None, 0, ctypes.byref(E))
    assert rc == 1
