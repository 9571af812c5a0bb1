"""ffa_sc_solve_noalloc (the SC solve on the library's static device array)
against ffa_sc_solve with a caller workspace, through the C ABI: identical
endpoint fields and flags on two calls in a row, the second with fewer
particles and a smaller grid, and a clear refusal of a problem larger than
the array."""
import ctypes

import pytest

from pyffa_amd import _lib
from tests import cases

pytestmark = pytest.mark.gpu


def _pair(n, grid, seed, precision):
    from pyffa_amd.FFAG_Bunch import FFAG_Bunch
    M = cases.bunch("stable", n=n, seed=seed, fi_spread=0.6)
    M[:, 7] = 1.0
    M[:, 8] = 1.0
    M[::53, 8] = 0.0
    out = []
    for _ in range(2):
        B = FFAG_Bunch(M.copy(), marcosize=1e13 / n, sc_grid_size=grid, BunchType="Boris", precision=precision)
        cases.set_bunch_seo(B)
        out.append(B)
    return out


def test_noalloc_solve_matches_the_workspace_solve(torch_cuda):
    torch = torch_cuda
    ctx = _lib.Context(0, torch.cuda.current_stream().cuda_stream)
    for n, grid, prec in ((300_000, (64, 128, 128, 4.0), "fp32"), (30_000, (32, 64, 32, 4.0), "fp64")):
        print(f"[case] noalloc solve n={n} grid={grid[:3]} {prec}", flush=True)
        A, B = _pair(n, grid, 7, prec)
        L = _lib.sc_layout(n, grid[0], grid[2], grid[1])
        assert L.total <= _lib.SC_NOALLOC_BYTES
        ws, _ = A.sc_workspace()
        ctx.check(ctx.lib.ffa_sc_solve(ctx.h, A.particles(), A.sc_params(first=False, s_last=0.5), ws.data_ptr()),
                  "ffa_sc_solve")
        ctx.check(ctx.lib.ffa_sc_solve_noalloc(ctx.h, B.particles(), B.sc_params(first=False, s_last=0.5)),
                  "ffa_sc_solve_noalloc")
        torch.cuda.synchronize()
        assert A.esc_prepre.abs().amax().item() > 0.0
        assert torch.equal(A.esc_prepre, B.esc_prepre) and torch.equal(A.esc_pre, B.esc_pre)
        assert torch.equal(A.flags, B.flags)
        del A, B, ws
    # a problem whose workspace exceeds the static array is refused before any launch
    A, _ = _pair(1000, (512, 256, 256, 4.0), 3, "fp32")
    p = A.particles()
    p.n = 10_000_000
    assert _lib.sc_layout(p.n, 512, 256, 256).total > _lib.SC_NOALLOC_BYTES
    from pyffa_amd.FFAG_Bunch import fixed_point_bits
    sp = A.sc_params(first=True, s_last=0.0)
    sp.bits = fixed_point_bits(p.n)
    rc = ctx.lib.ffa_sc_solve_noalloc(ctx.h, ctypes.byref(p), sp)
    assert rc == -1 and b"static array" in ctx.lib.ffa_last_error(ctx.h)
