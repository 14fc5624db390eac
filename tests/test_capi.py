"""CPU checks of the C ABI boundary: the in-tree library loads and exports
every entry point include/philip_dg.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

import philip_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "philip_dg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dgx_[a-z0-9_]+)\s*\(", src)))


def test_header_lists_entry_points():
    syms = declared_symbols()
    assert "dgx_assemble_residual" in syms and "dgx_step_in_time" in syms
    assert set(philip_amd.EXPORTED) <= set(syms)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(philip_amd.LIB_PATH):
        pytest.fail("libphilip_dg.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(philip_amd.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_default_params_without_gpu():
    p = philip_amd.params()
    assert p.dim == 3 and p.poly_degree == 2 and abs(p.gamma_gas - 1.4) < 1e-15
    assert p.conv_flux == philip_amd.LXF and p.two_point_flux == philip_amd.KG
