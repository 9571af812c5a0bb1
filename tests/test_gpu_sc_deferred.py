"""The FP32 tier with each SC solve's gather and endpoint roll run inside the
push launch that follows (ffa_sc_solve_field + ffa_push_boris_fused_f32_sc,
LocalParameters defer_sc_gather, the tracking loop's default) against the
solve's own gather kernel (ffa_sc_solve + ffa_push_boris_fused_f32):
bit-identical state, flags, RF phase and SC endpoints after steps with space
charge, RF, the aperture and injection between solves, for the gather from
the field cubes and the gather that differentiates phi."""
import ctypes

import numpy as np
import pytest
import torch

from pyffa_amd import _lib
from pyffa_amd import workloads as W

pytestmark = pytest.mark.gpu

KEYS = ("r", "vr", "z", "vz", "fi", "fidot", "rf_phase", "flags", "esc_pre", "esc_prepre")


def _bunch(M, cubes):
    from pyffa_amd.FFAG_Bunch import FFAG_Bunch
    n = M.shape[0]
    B = FFAG_Bunch(M.copy(), marcosize=1e13 / n, sc_grid_size=(32, 64, 64, 4.0), BunchType="Boris",
                   precision="fp32")
    W.set_bunch_seo(B)
    B.sc_field_cubes = cubes
    return B


def _run(defer, M, gp, nsteps, dt, cubes, sort_every):
    from pyffa_amd.FFAG_track import FFAG_RungeKutta
    B = _bunch(M, cubes)
    lp = dict(stop_condition=dict(max_stepsN=nsteps, max_turn=1e9, max_time=1.0), time_step=dt, step_dumps=None,
              enable_SC=True, SC_type=1, SC_step=20, precision="fp32", verbose=False, sort_every=sort_every,
              defer_sc_gather=defer)
    FFAG_RungeKutta().rk4_solve_dt_bunch3_withSC_boris(None, 0.0, B, gp, lp)
    assert B._pending_gather is None
    return {k: B.ordered(k) for k in KEYS}


@pytest.mark.parametrize("cubes,sort_every", [(False, 100), (True, 0)])
def test_push_with_the_solve_gather_is_bit_identical(torch_cuda, tmp_path, cubes, sort_every):
    print(f"[case] deferred SC gather vs the solve's gather, cubes={cubes}, sort_every={sort_every}", flush=True)
    gp = W.global_parameters("stable", max_order=2, rf=dict(v0=100e3), aperture=W.APERTURE, tmpdir=str(tmp_path))
    n, dt, nsteps = 20000, 0.05e-9, 260
    M = W.bunch("stable", n=n, seed=13)
    # injection times (column 6, ns) staggered over the run: solves see
    # growing bunches, blocks carry injection ranges after their first step
    M[:, 6] = np.sort(np.random.default_rng(6).uniform(0.0, nsteps * dt * 0.9e9, n))
    a = _run(False, M, gp, nsteps, dt, cubes, sort_every)
    b = _run(True, M, gp, nsteps, dt, cubes, sort_every)
    alive = int(((a["flags"] & _lib.F_ALIVE) != 0).sum())
    injected = int(((a["flags"] & _lib.F_INJ) != 0).sum())
    sc_both = int(((a["flags"] & _lib.F_SCPRE) != 0).sum())
    print(f"[case] injected {injected}, alive {alive}, with an SC endpoint {sc_both} of {n}", flush=True)
    assert injected == n and 0.5 * n < alive and sc_both > 0
    for k in a:
        assert torch.equal(a[k], b[k]), k


def test_gather_alone_matches_the_solve(torch_cuda, tmp_path):
    """ffa_push_boris_fused_f32_sc with no steps is the solve's gather and
    roll: the endpoint fields and flags equal ffa_sc_solve's, on the first
    solve and on a later one (both endpoints present)."""
    from pyffa_amd.FFAG_track import FFAG_RungeKutta
    print("[case] deferred gather with nsteps = 0 vs ffa_sc_solve", flush=True)
    gp = W.global_parameters("stable", max_order=2, rf=dict(v0=100e3), aperture=W.APERTURE, tmpdir=str(tmp_path))
    n = 5000
    M = W.bunch("stable", n=n, seed=4)
    out = []
    for defer in (False, True):
        B = _bunch(M, True)
        rk = FFAG_RungeKutta()
        # 40 steps through the loop (two solves), then a third solve by hand
        lp = dict(stop_condition=dict(max_stepsN=40, max_turn=1e9, max_time=1.0), time_step=0.05e-9,
                  step_dumps=None, enable_SC=True, SC_type=1, SC_step=20, precision="fp32", verbose=False,
                  defer_sc_gather=False)
        rk.rk4_solve_dt_bunch3_withSC_boris(None, 0.0, B, gp, lp)
        ctx = rk.ctx
        B.sc_solve_device(ctx, first=False, s_last=0.95, defer_gather=defer)
        if defer:
            steps = (_lib.FfaStep * 1)()
            inj = (_lib._i64 * 2)()
            B.push_fused_f32(ctx, steps, inj, 0)
        torch.cuda.synchronize()
        out.append({k: B.ordered(k) for k in ("flags", "esc_pre", "esc_prepre", "r", "z")})
    both = int(((out[0]["flags"] & _lib.F_SCPREPRE) != 0).sum())
    assert both > 0.5 * n
    for k in out[0]:
        assert torch.equal(out[0][k], out[1][k]), k


def test_solve_field_refuses_bad_arguments(torch_cuda):
    """The new entry points validate like ffa_sc_solve (NULL workspace)."""
    ctx = _lib.Context(0, None)
    sp = _lib.FfaScParams()
    p = _lib.FfaParticles()
    rc = ctx.lib.ffa_sc_solve_field(ctx.h, ctypes.byref(p), ctypes.byref(sp), None)
    assert rc != 0
    steps = (_lib.FfaStep * 1)()
    inj = (_lib._i64 * 2)()
    rc = ctx.lib.ffa_push_boris_fused_f32_sc(ctx.h, ctypes.byref(p), steps, inj, 0, ctypes.byref(sp), None)
    assert rc != 0
