"""The C-ABI library loads and exports every symbol include/ffa_hip.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

from pyffa_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "ffa_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return set(re.findall(r"^\s*(?:int|const char \*)\s*(ffa_\w+)\s*\(", src, flags=re.M))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    decl = declared_symbols()
    assert decl, "no declarations parsed"
    assert decl == set(_lib.EXPORTED), decl ^ set(_lib.EXPORTED)
    for name in decl:
        assert hasattr(lib, name), name


def test_struct_sizes_match_header_layout():
    # ffa_step: 5 doubles + 3 + 3
    assert ctypes.sizeof(_lib.FfaStep) == 11 * 8
    assert ctypes.sizeof(_lib.FfaParticles) == 10 * 8 + 8 + 8
    assert ctypes.sizeof(_lib.FfaScLayout) == 16 * 8
    assert ctypes.sizeof(_lib.FfaScResult) == 32 * 8


def test_workspace_layout_is_aligned_and_ordered():
    L = _lib.sc_layout(1000, 33, 64, 32)
    offs = [L.grid, L.cell_count, L.cbuf, L.phi, L.ex, L.ez, L.px, L.ppr0, L.pslice, L.slice_count, L.partials,
            L.result, L.records, L.slice_start, L.cursor]
    assert all(o % 256 == 0 for o in offs)
    assert offs == sorted(offs) and L.total > offs[-1]
    assert L.cell_count - L.grid >= 8 * 33 * 64 * 32


def test_invalid_grid_is_rejected_without_gpu():
    lib = _lib.load()
    L = _lib.FfaScLayout()
    rc = lib.ffa_sc_workspace_layout(100, 16, 48, 64, ctypes.byref(L))
    assert rc == -1
    assert b"powers of two" in lib.ffa_last_error(None)


def test_missing_gpu_fails_loudly():
    import torch
    import numpy as np
    import pytest
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from pyffa_amd.FFAG_Bunch import FFAG_Bunch
    with pytest.raises(_lib.FfaError):
        FFAG_Bunch(np.zeros((4, 16)))
