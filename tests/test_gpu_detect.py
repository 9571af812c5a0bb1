"""GPU tests of the position detectors: the crossing kernel against the
oracle (bit-exact index sets, boundary and NaN cases included), the bunch's
crossing query in the input order after an azimuth sort, and the detector
files of a tracking run against a step-by-step host computation."""
import os

import numpy as np
import pytest

from oracle import detect as D
from tests import cases

pytestmark = pytest.mark.gpu

TWO_PI = 2 * np.pi


def test_kernel_matches_oracle(torch_cuda):
    torch = torch_cuda
    from pyffa_amd import _lib
    ctx = _lib.Context(0, torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(7)
    n = 300001
    pre = rng.uniform(-20.0, 40.0, n)
    post = pre + rng.uniform(-0.5, 7.0, n)          # forward, backward and multi-wrap steps
    alive = rng.random(n) < 0.9
    pre[::1001] = np.nan
    post[5::2003] = np.nan
    targets = [0.0, 1.0, np.pi, float(D.reduce_azimuth(pre[10])), float(D.reduce_azimuth(post[20])),
               float(np.nextafter(TWO_PI, 0.0))]
    post[30] = pre[30] + TWO_PI                      # the same azimuth one turn later
    fl = np.where(alive, _lib.F_ALIVE | _lib.F_INJ, _lib.F_INJ).astype(np.uint8)
    d_pre = torch.tensor(pre, device="cuda")
    d_post = torch.tensor(post, device="cuda")
    d_fl = torch.tensor(fl, device="cuda")
    idx = torch.empty(n, dtype=torch.int64, device="cuda")
    count = torch.empty(1, dtype=torch.int64, device="cuda")
    for tgt in targets:
        for m in (n, 777, 0):                        # later calls are smaller: the counter restarts
            print(f"[detect] target={tgt!r} n={m}", flush=True)
            ctx.check(ctx.lib.ffa_detect_crossing(ctx.h, _lib._vp(d_pre.data_ptr()), _lib._vp(d_fl.data_ptr()),
                                                  _lib._vp(d_post.data_ptr()), m, tgt, _lib._vp(idx.data_ptr()),
                                                  _lib._vp(count.data_ptr())), "ffa_detect_crossing")
            k = int(count.item())
            got = np.sort(idx[:k].cpu().numpy())
            want = np.flatnonzero(D.crossing_mask(pre[:m], post[:m], alive[:m], tgt))
            assert np.array_equal(got, want)
            if m == n:
                assert 0.01 * n < k < 0.9 * n
    # a target outside [0, 2 pi) is refused
    rc = ctx.lib.ffa_detect_crossing(ctx.h, _lib._vp(d_pre.data_ptr()), _lib._vp(d_fl.data_ptr()),
                                     _lib._vp(d_post.data_ptr()), n, TWO_PI, _lib._vp(idx.data_ptr()),
                                     _lib._vp(count.data_ptr()))
    assert rc != 0


def test_bunch_crossing_query_after_sort(torch_cuda):
    """detect_crossing reports input rows (increasing) and the device
    positions of the same particles, with the bunch in azimuth order."""
    torch = torch_cuda
    from pyffa_amd.FFAG_Bunch import FFAG_Bunch
    n = 20000
    rng = np.random.default_rng(2)
    M = cases.bunch("stable", n=n, seed=3)
    M[:, 4] = rng.uniform(-10.0, 30.0, n)
    M[:, 7] = 1
    B = FFAG_Bunch(M, precision="fp32")
    B._uninject_head = n
    fl = B.flags.cpu().numpy()
    fl[rng.random(n) < 0.1] &= ~np.uint8(2)           # some lost particles
    B.flags.copy_(torch.tensor(fl, device="cuda"))
    print("[detect] sort, record, advance, query", flush=True)
    B.sort_by_azimuth()
    B.UpdatePreSteps()
    pre = B.ordered("fi").cpu().numpy()
    alive = (B.ordered("flags").cpu().numpy() & 2) != 0
    dphi = torch.tensor(rng.uniform(0.0, 0.4, n), device="cuda", dtype=torch.float64)
    B.fi.add_(dphi[B.pid])                              # per input row, applied in device order
    post = pre + dphi.cpu().numpy()
    assert np.array_equal(B.ordered("fi").cpu().numpy(), post)
    for deg in (0.0, 45.0, 359.0):
        tgt = np.deg2rad(deg) % TWO_PI
        rows, pos = B.detect_crossing(tgt)
        want = np.flatnonzero(D.crossing_mask(pre, post, alive, tgt))
        assert np.array_equal(rows, want)
        assert np.array_equal(B.pid[pos].cpu().numpy(), want)
        sub = np.sort(rng.choice(n, 500, replace=False))
        mask = B.detect_crossing(tgt, sub)
        assert np.array_equal(mask, D.crossing_mask(pre[sub], post[sub], alive[sub], tgt))
        assert np.array_equal(B.rows(sub[mask]), B.LocalBunch[sub[mask]])


def _dumps(folder, gids):
    from pyffa_amd import FFAG_dump as FD
    d = FD.Dumps()
    d.add_dump(FD.StepDump(0.5e-9, 10e-9, 1e-9, gids[:8], save_folder=os.path.join(folder, "step")))
    d.add_dump(FD.StepDumpBunch(2e-9, 8e-9, 2e-9, save_folder=os.path.join(folder, "stepbunch")))
    d.add_dump(FD.PositionDump(30.0, 0.0, 13e-9, gids[::3], save_folder=os.path.join(folder, "pos")))
    d.add_dump(FD.PositionDumpBunch(200.0, 1e-9, 12e-9, save_folder=os.path.join(folder, "posbunch")))
    return d


def _load(folder):
    out = {}
    for root, _, files in os.walk(folder):
        for f in files:
            with np.load(os.path.join(root, f)) as z:
                out[os.path.relpath(os.path.join(root, f), folder)] = z["particles"]
    return out


def _lp(steps, dumps, **kw):
    lp = dict(stop_condition=dict(max_stepsN=steps, max_turn=1e9, max_time=1.0), time_step=0.05e-9,
              step_dumps=dumps, enable_SC=False, SC_step=20, verbose=False)
    lp.update(kw)
    return lp


STEPS = 300


def test_tracker_detectors_match_step_by_step(torch_cuda, tmp_path):
    """FP64 Boris run with all four detectors (multi-step launches between
    detector events) against the same run advanced one step per call with
    the detector records computed on the host from LocalBunch snapshots."""
    from pyffa_amd.FFAG_Bunch import FFAG_Bunch
    from pyffa_amd.FFAG_ParasAndConversion import FFAG_ConversionTools
    from pyffa_amd.FFAG_track import FFAG_RungeKutta
    conv = FFAG_ConversionTools()
    _, gp = cases.machines("stable", max_order=2, rf=dict(v0=100e3), aperture=cases.APERTURE, tmpdir=str(tmp_path))
    n = 3000
    M = cases.bunch("stable", n=n, seed=21)
    M[:, 15] = 5000 + np.arange(n)[::-1]
    gids = M[:, 15].astype(np.int64)
    print("[detect] tracking with detectors", flush=True)
    B = FFAG_Bunch(M, BunchType="Boris")
    tr = FFAG_RungeKutta()
    tr.rk4_solve_dt_bunch3_withSC_boris(None, 0.0, B, gp, _lp(STEPS, _dumps(str(tmp_path / "run"), gids)))
    got = _load(str(tmp_path / "run"))
    final = B.LocalBunch.copy()

    print("[detect] step-by-step reference", flush=True)
    B2 = FFAG_Bunch(M, BunchType="Boris")
    tr2 = FFAG_RungeKutta()
    t = 0.0
    tgt_pb, tgt_p = np.deg2rad(200.0) % TWO_PI, np.deg2rad(30.0) % TWO_PI
    pb, pos, step, sb = [], {}, {}, {}
    for k in range(STEPS):
        tr2.rk4_solve_dt_bunch3_withSC_boris(None, t, B2, gp, _lp(k + 1, None, start_step=k))
        if k == 0:
            pre = M.copy()
            pre[:, 7] = 1                               # injected at the first step
        post = B2.LocalBunch.copy()
        tn = tr2.last_run["t"]
        assert tn == t + 0.05e-9
        alive = pre[:, 8] == 1
        if 1e-9 < tn < 12e-9:
            h = np.flatnonzero(D.crossing_mask(pre[:, 4], post[:, 4], alive, tgt_pb))
            c = conv.ConvertVrzek2Przek_boris(post[h])
            c[:, 9] = np.mod(c[:, 9], TWO_PI)
            pb.append(np.hstack((c, np.full((h.size, 1), tn))))
        if 0.0 < tn < 13e-9:
            rows = np.arange(0, n, 3)
            h = rows[D.crossing_mask(pre[rows, 4], post[rows, 4], alive[rows], tgt_p)]
            for r in h:
                c = conv.ConvertVrzek2Przek_boris(post[r:r + 1])
                c[:, 9] = np.mod(c[:, 9], TWO_PI)
                pos.setdefault(int(gids[r]), []).append(np.hstack((c, [[tn]]))[0])
        if 0.5e-9 < t < 10e-9 and 0.5e-9 < tn < 10e-9 and \
                np.floor((t - 0.5e-9) / 1e-9 + 1e-12) != np.floor((tn - 0.5e-9) / 1e-9 + 1e-12):
            for r in range(8):
                if alive[r]:
                    c = conv.ConvertVrzek2Przek_boris(post[r:r + 1])
                    step.setdefault(int(gids[r]), []).append(np.hstack((c, [[tn]]))[0])
        if 2e-9 < tn < 8e-9 and (t - 2e-9) // 2e-9 != (tn - 2e-9) // 2e-9:
            live = post[post[:, 8] == 1]
            sb[f"stepbunch/time_{tn * 1e9:.2f}_rank_0.npz"] = np.hstack(
                (conv.ConvertVrzek2Przek_boris(live), np.full((live.shape[0], 1), tn)))
        pre = post
        t = tn
    assert np.array_equal(final, B2.LocalBunch, equal_nan=True)

    a = got.pop("posbunch/crossing_angle_200.0_start_1.000_end_12.000_rank_0.npz")
    want = np.vstack(pb)
    print(f"[detect] {want.shape[0]} crossings at 200 deg", flush=True)
    assert want.shape[0] > 20
    assert np.array_equal(a, want, equal_nan=True)
    assert len(pos) > 5
    for gid, rows in pos.items():
        assert np.array_equal(got.pop(f"pos/particle_{gid}_angle_30.0_rank_0.npz"), np.vstack(rows), equal_nan=True)
    assert len(step) >= 6
    for gid, rows in step.items():
        assert np.array_equal(got.pop(f"step/particle_{gid}_rank_0.npz"), np.vstack(rows), equal_nan=True)
    assert len(sb) >= 2
    for name, rows in sb.items():
        assert np.array_equal(got.pop(name), rows, equal_nan=True)
    assert not got, sorted(got)


def test_fp32_detectors_independent_of_sort(torch_cuda, tmp_path):
    """FP32 tier without SC: particles pushed in azimuth order (sorted every
    100 steps) or in input order give the same detector files."""
    from pyffa_amd.FFAG_Bunch import FFAG_Bunch
    from pyffa_amd.FFAG_track import FFAG_RungeKutta
    _, gp = cases.machines("stable", max_order=2, rf=dict(v0=100e3), aperture=cases.APERTURE, tmpdir=str(tmp_path))
    n = 3000
    M = cases.bunch("stable", n=n, seed=22)
    gids = M[:, 15].astype(np.int64)
    out = []
    for sort_every in (100, 0):
        print(f"[detect] fp32 sort_every={sort_every}", flush=True)
        folder = str(tmp_path / f"s{sort_every}")
        B = FFAG_Bunch(M, BunchType="Boris", precision="fp32")
        FFAG_RungeKutta().rk4_solve_dt_bunch3_withSC_boris(
            None, 0.0, B, gp, _lp(STEPS, _dumps(folder, gids), precision="fp32", sort_every=sort_every,
                                  max_block=100))
        out.append(_load(folder))
    a, b = out
    assert sorted(a) == sorted(b) and len(a) > 10
    assert a["posbunch/crossing_angle_200.0_start_1.000_end_12.000_rank_0.npz"].shape[0] > 20
    for k in a:
        assert np.array_equal(a[k], b[k], equal_nan=True), k
