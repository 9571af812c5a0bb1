"""The FP64 Boris path with its steps fused into one launch per block
(ffa_push_boris_f64_fused, LocalParameters fuse_fp64) against the same path
with one launch per step: bit-identical state, flags, RF phase and SC
endpoints after steps with space charge, RF, the aperture and injection
between solves."""
import numpy as np
import pytest
import torch

from pyffa_amd import _lib
from pyffa_amd import workloads as W

pytestmark = pytest.mark.gpu


def _run(fuse, M, gp, nsteps, dt, sort_every):
    from pyffa_amd.FFAG_Bunch import FFAG_Bunch
    from pyffa_amd.FFAG_track import FFAG_RungeKutta
    n = M.shape[0]
    B = FFAG_Bunch(M.copy(), marcosize=1e13 / n, sc_grid_size=(32, 64, 64, 4.0), BunchType="Boris")
    W.set_bunch_seo(B)
    lp = dict(stop_condition=dict(max_stepsN=nsteps, max_turn=1e9, max_time=1.0), time_step=dt, step_dumps=None,
              enable_SC=True, SC_type=1, SC_step=20, verbose=False, sort_every=sort_every, fuse_fp64=fuse)
    FFAG_RungeKutta().rk4_solve_dt_bunch3_withSC_boris(None, 0.0, B, gp, lp)
    return {k: B.ordered(k) for k in ("r", "vr", "z", "vz", "fi", "fidot", "rf_phase", "flags", "esc_pre",
                                      "esc_prepre")}


@pytest.mark.parametrize("sort_every", [0, 100])
def test_fused_fp64_steps_are_bit_identical(torch_cuda, tmp_path, sort_every):
    print(f"[case] fused fp64 steps vs one launch per step, sort_every={sort_every}", flush=True)
    gp = W.global_parameters("stable", max_order=2, rf=dict(v0=100e3), aperture=W.APERTURE, tmpdir=str(tmp_path))
    n, dt, nsteps = 20000, 0.05e-9, 260
    M = W.bunch("stable", n=n, seed=7)
    # injection times (column 6, ns) staggered over the run: blocks of steps
    # carry injection ranges after their first step
    M[:, 6] = np.sort(np.random.default_rng(5).uniform(0.0, nsteps * dt * 0.9e9, n))
    a = _run(False, M, gp, nsteps, dt, sort_every)
    b = _run(True, M, gp, nsteps, dt, sort_every)
    alive = int(((a["flags"] & _lib.F_ALIVE) != 0).sum())
    injected = int(((a["flags"] & _lib.F_INJ) != 0).sum())
    print(f"[case] injected {injected}, alive {alive} of {n}", flush=True)
    assert injected == n and 0.5 * n < alive
    for k in a:
        assert torch.equal(a[k], b[k]), k
