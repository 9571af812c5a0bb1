"""The C-ABI library loads and exports every symbol include/ffa_hip.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

import pytest

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
    assert ctypes.sizeof(_lib.FfaScLayout) == 18 * 8
    assert ctypes.sizeof(_lib.FfaScResult) == 32 * 8
    assert ctypes.sizeof(_lib.FfaMlpDesc) == 8 * 8 + 3 * 32 * 4 + 8
    assert ctypes.sizeof(_lib.FfaAdam) == 6 * 4
    assert ctypes.sizeof(_lib.FfaMlpLayout) == 7 * 8


def test_header_constants_match_binding():
    src = open(os.path.join(ROOT, "include", "ffa_hip.h")).read()
    for name, val in (("FFA_MLP_NFREQ", _lib.MLP_NFREQ), ("FFA_MLP_NPARAM", _lib.MLP_NPARAM),
                      ("FFA_MLP_PACKED_BYTES", _lib.MLP_PACKED_BYTES), ("FFA_MLP_KCHUNK", _lib.MLP_KCHUNK)):
        m = re.search(r"#define\s+%s\s+(\d+)" % name, src)
        assert m and int(m.group(1)) == val, name


def test_mlp_workspace_layout():
    L = _lib.mlp_layout(1000)
    assert L.batch_padded == 1024
    offs = [L.xfm, L.dzfm, L.zs, L.partials, L.loss]
    assert all(o % 256 == 0 for o in offs) and offs == sorted(offs) and L.total > offs[-1]
    assert L.dzfm - L.xfm >= (64 + 4 * 128) * 1024 * 2
    assert L.loss - L.partials >= 2 * _lib.MLP_NPARAM * 4
    lib = _lib.load()
    assert lib.ffa_mlp_workspace_layout(0, ctypes.byref(L)) == -1


def test_workspace_layout_is_aligned_and_ordered():
    L = _lib.sc_layout(1000, 33, 64, 32)
    offs = [L.grid, L.cell_count, L.cbuf, L.phi, L.exz, L.px, L.ppr0, L.pslice, L.slice_count, L.partials,
            L.result, L.records, L.slice_start, L.block_offset, L.records_by_row, L.row_start, L.tables]
    assert all(o % 256 == 0 for o in offs)
    assert offs == sorted(offs) and L.total > offs[-1]
    assert L.cell_count - L.grid >= 8 * 33 * 64 * 32
    assert L.records_by_row - L.block_offset >= 4 * _lib.FFA_SC_BLOCKS * 33


def test_invalid_grid_is_rejected_without_gpu():
    lib = _lib.load()
    L = _lib.FfaScLayout()
    # 2-3-5 smooth sizes in [32, 1024] are accepted (160 = 2^5 5, 180 = 2^2 3^2 5, 48 = 2^4 3)
    for nz, nx in ((160, 160), (180, 160), (48, 64), (1024, 1024)):
        assert lib.ffa_sc_workspace_layout(100, 16, nz, nx, ctypes.byref(L)) == 0
    # a prime factor above 5, or a size outside [32, 1024], is refused
    for nz, nx in ((112, 64), (64, 77), (16, 64), (64, 2048)):
        rc = lib.ffa_sc_workspace_layout(100, 16, nz, nx, ctypes.byref(L))
        assert rc == -1
        assert b"no prime factor above 5" in lib.ffa_last_error(None)


def test_missing_gpu_fails_loudly():
    import torch
    import numpy as np
    import pytest
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from pyffa_amd.FFAG_Bunch import FFAG_Bunch
    with pytest.raises(_lib.FfaError):
        FFAG_Bunch(np.zeros((4, 16)))


STRUCTS = [("FfaParticles", "ffa_particles"), ("FfaMachine", "ffa_machine"), ("FfaStep", "ffa_step"),
           ("FfaScParams", "ffa_sc_params"), ("FfaScResult", "ffa_sc_result"), ("FfaScLayout", "ffa_sc_layout"),
           ("FfaMlpDesc", "ffa_mlp_desc"), ("FfaAdam", "ffa_adam"), ("FfaMlpLayout", "ffa_mlp_layout"),
           ("FfaTaylorGen", "ffa_taylor_gen"), ("FfaOrbit", "ffa_orbit"),
           ("FfaSpiralGap", "ffa_spiral_gap")]


def test_ctypes_mirrors_match_the_compiled_header(tmp_path):
    """Every ctypes struct has the size and field offsets that a C compiler
    gives the header's struct (gcc on include/ffa_hip.h)."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    inc = os.path.join(ROOT, "include")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "ffa_hip.h"', "int main(void) {"]
    want = []
    for py, c in STRUCTS:
        cls = getattr(_lib, py)
        lines.append(f'printf("%zu\\n", sizeof({c}));')
        want.append(ctypes.sizeof(cls))
        for f in cls._fields_:
            lines.append(f'printf("%zu\\n", offsetof({c}, {f[0]}));')
            want.append(getattr(cls, f[0]).offset)
    lines += ["return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run([gcc, "-std=c11", "-I", inc, str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == want
