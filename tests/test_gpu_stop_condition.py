"""The tracking loop's stop condition (src/FFAG_track.py:559-577): every
1000 steps the mean azimuth over all particles with a non-NaN azimuth
(uninjected and lost ones included) is compared with 2 pi max_turn, and the
time with max_time; the loop ends at the first 1000-step boundary that
passes either.  The GPU loop (ffa_reduce_phi) must stop at the same step as
the oracle's restatement, and the boundary values must bracket the limit."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import cases

pytestmark = pytest.mark.gpu
DT = 0.05e-9


def _bunch(n):
    M = cases.bunch("stable", n=n, seed=8)
    M[:, 6] = 0.0                       # injected at the first step ...
    M[-40:, 6] = 1e6                    # ... but the last 40 never are
    M[-30:, 4] = np.nan                 # 30 of those without an azimuth: excluded from the mean
    return M


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_max_turn_stops_at_the_1000_step_boundary(torch_cuda, precision, tmp_path):
    from pyffa_amd.FFAG_Bunch import FFAG_Bunch
    from pyffa_amd.FFAG_track import FFAG_RungeKutta
    print(f"[case] stop condition max_turn {precision}", flush=True)
    mc, gp = cases.machines("stable", max_order=2)
    n = 4096
    M = _bunch(n)
    # the mean azimuth at the 1000-step boundaries, from an unstopped oracle run
    tr = O.Tracker(mc, M.copy())
    means = []
    for b in range(1, 6):
        tr.run(0.0 if b == 1 else tr.t, DT, 1000 * b, start_step=1000 * (b - 1))
        fi = tr.M[:, 4]
        means.append(float(np.mean(fi[~np.isnan(fi)])))
    # a limit between the boundary means at 2000 and 3000 steps
    max_turn = 0.5 * (means[1] + means[2]) / (2 * np.pi)
    assert means[1] <= 2 * np.pi * max_turn < means[2]
    ref = O.Tracker(mc, M.copy())
    stop_ref = ref.run(0.0, DT, 5000, max_turn=max_turn, max_time=1.0)
    assert stop_ref == 3000
    B = FFAG_Bunch(M.copy(), marcosize=1.0, sc_grid_size=(16, 32, 32, 4.0), BunchType="Boris", precision=precision)
    lp = dict(stop_condition=dict(max_stepsN=5000, max_turn=max_turn, max_time=1.0), time_step=DT, step_dumps=None,
              enable_SC=False, SC_type=1, SC_step=20, precision=precision, verbose=False)
    rk = FFAG_RungeKutta()
    rk.rk4_solve_dt_bunch3_withSC_boris(None, 0.0, B, gp, lp)
    assert rk.last_run["steps"] == 3000
    fi = B.LocalBunch[:, 4]
    assert np.isnan(fi[-30:]).all() and not np.isnan(fi[:-30]).any()
    assert np.mean(fi[~np.isnan(fi)]) > 2 * np.pi * max_turn


def test_max_time_stops_at_the_next_boundary(torch_cuda):
    from pyffa_amd.FFAG_Bunch import FFAG_Bunch
    from pyffa_amd.FFAG_track import FFAG_RungeKutta
    print("[case] stop condition max_time", flush=True)
    _, gp = cases.machines("stable", max_order=2)
    B = FFAG_Bunch(_bunch(1024), marcosize=1.0, sc_grid_size=(16, 32, 32, 4.0), BunchType="Boris")
    # t passes max_time during step 1200: the check at 2000 steps ends the run
    lp = dict(stop_condition=dict(max_stepsN=5000, max_turn=1e9, max_time=1200.5 * DT), time_step=DT,
              step_dumps=None, enable_SC=False, SC_type=1, SC_step=20, verbose=False)
    rk = FFAG_RungeKutta()
    rk.rk4_solve_dt_bunch3_withSC_boris(None, 0.0, B, gp, lp)
    assert rk.last_run["steps"] == 2000
    assert rk.last_run["t"] == pytest.approx(2000 * DT, rel=1e-12)
