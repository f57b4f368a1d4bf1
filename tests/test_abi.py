"""CPU checks of the boundary: the HIP library loads (no GPU needed to dlopen)
and exports every symbol include/gpt_amd_wc.h declares; the host-side
parameter handling mirrors gpt's (lib/gpt/qcd/fermion/wilson.py:63-96)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "gpt_amd_wc.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(wc_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from gpt_amd import _lib

    lib = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 19
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # and the ctypes table covers the whole header
    assert sorted(_lib.SIGNATURES) == syms


def test_params_struct_layout_matches_header():
    from gpt_amd import _lib

    # int[4], 6 doubles, int, double[4][2], 2 ints (+ padding to 8)
    assert ctypes.sizeof(_lib.wc_params) == 16 + 6 * 8 + 8 + 64 + 8
    assert _lib.wc_params.boundary_phases.offset == 16 + 48 + 8


def test_error_paths_without_gpu():
    """argument validation happens before any device call"""
    from gpt_amd import _lib

    lib = _lib.load()
    p = _lib.wc_params()
    p.fdims[:] = [4, 4, 3, 8]  # odd dimension
    p.xi_0 = p.nu = p.cF = 1.0
    h = ctypes.c_void_p()
    rc = lib.wc_create(ctypes.byref(p), ctypes.c_void_p(1), None, ctypes.byref(h))
    assert rc == _lib.WC_ERR_ARG
    assert b"even" in lib.wc_last_error()
    rc = lib.wc_axpy(7, 1, 1.0, 0.0, None, None, None, None)
    assert rc == _lib.WC_ERR_ARG and b"precision" in lib.wc_last_error()
    rc = lib.wc_apply(None, 2001, None, None, 0, None)
    assert rc == _lib.WC_ERR_ARG


def test_grid_validation():
    import gpt_amd as g

    with pytest.raises(ValueError):
        g.grid([4, 4, 4, 5], g.double)
    gr = g.grid([4, 4, 4, 8], g.double)
    assert gr.Vh == 256 and gr.checkerboarded(g.redblack).cb is g.redblack
