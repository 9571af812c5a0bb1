"""The main_track entry on the GPU: a config_track.json job (painted
injection, RF, aperture, SC, detectors) against the oracle tracker started
from the same bunch, and a restart from a StepDumpBunch file against the
uninterrupted run."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import cases

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DT = 0.05e-9


def _setup(tmp, *, enable_sc, steps, dumps):
    from pyffa_amd import lattice as Lt
    lat = cases.lattice("stable", rstep=0.005, theta_step_deg=0.02)
    bmap = os.path.join(tmp, "Bmap")
    Lt.write_bmap(bmap, lat, 2)
    seo = cases.seo("stable")
    seo_dir = os.path.join(bmap, "resultsSEO")
    Lt.write_seo(seo_dir, seo)
    Lt.write_twiss(seo_dir, seo, 2.2, 2.9, 0.1, -0.2)
    emap = cases.emap_file(lat, tmp)
    curve = os.path.join(tmp, "curve.paint")
    np.savetxt(curve, np.array([[0.0, 0.0, 0.0, 0.0, 0.0], [20.0, 2e-3, 2e-4, 1e-3, 1e-4]]), header="t x px y py")
    cfg = {
        "config": {"date": 0.0, "MPI_process_num": 1, "thread_num": 1},
        "track": {"start_EkMeV": 300.0, "time_start": -200e-9, "time_step": DT, "start_azimuth": 15.0,
                  "stop_condition": {"max_stepsN": steps, "max_turn": 1e9, "max_time": 1.0}, "solver_type": 1},
        "SC": {"enable_SC": enable_sc, "SC_type": 1, "grid_size_rfz": [16, 32, 32, 4.0], "step": 20},
        "BmapAndSEO": {"maps": bmap, "max_order": 2, "SEO_FileName": "resultsSEO"},
        "Emap": {"maps": emap, "enable": True},
        "BunchPara": {"ParticleDensity": 1e9, "ParticleNum": 400, "TransverseREmit": 10.0, "TransverseZEmit": 10.0,
                      "LongitudeT": 20.0, "LongitudeDEk": 1.0, "InjTimeNanoSec": 0, "TransverseDistType": "gauss",
                      "LongitudeDistType": "UniformGauss"},
        "Paint": {"enable": True, "MaxBunchNum": 3, "TimeInterval": 5.0, "Curve": curve},
        "Aperture": {"enable": True, "Rmin": cases.APERTURE[0], "Rmax": cases.APERTURE[1],
                     "Zmin": cases.APERTURE[2], "Zmax": cases.APERTURE[3]},
        "DumpPara": {"modules": dumps},
        "check": {"CheckBunch": False, "PlotRMS": [4, 4, 4]},
    }
    path = os.path.join(tmp, "config_track.json")
    with open(path, "w") as f:
        json.dump(cfg, f, indent=1)
    return path, cfg


def test_main_track_job_matches_the_oracle(tmp_path):
    from pyffa_amd import FFAG_Distribution as FD
    from pyffa_amd import main_track
    print("main_track job vs oracle", flush=True)
    steps = 5000                                       # -200 ns .. +50 ns: all three sub-bunches injected
    dumps = [{"type": "PositionDumpBunch", "start_time": -10e-9, "end_time": 10e-9, "dump_azimuth": [20.0],
              "save_folder": str(tmp_path / "out" / "tbt")},
             {"type": "StepDumpBunch", "start_time": 0.0, "end_time": 60e-9, "interval_time": 20e-9,
              "save_folder": str(tmp_path / "out" / "snap"), "filtter_survive_paticles": True}]
    path, cfg = _setup(str(tmp_path), enable_sc=True, steps=steps, dumps=dumps)
    # 1e6 protons per macro-particle: the GPU's complex64 Poisson solve
    # differs from the oracle's FP64 one by ~1e-7 of the field, which a
    # strong space-charge force (1e9 per macro-particle) amplifies over the
    # 1200 steps after injection to ~1e-3 in z; at this intensity the free
    # run stays close enough to check the whole job against the oracle
    cfg["BunchPara"]["ParticleDensity"] = 1e6
    with open(path, "w") as f:
        json.dump(cfg, f)
    FD.set_seed(11)
    rk, B, gp = main_track.main(path)
    assert rk.last_run["steps"] == steps
    assert B.InjectedNum_Local == 1200 and B.TotalParticleNum == 1200
    ref = O.Tracker(cases.oracle_machine(gp), B.StaticBunchLocal.copy(), enable_sc=True, sc_step=20,
                    grid=tuple(cfg["SC"]["grid_size_rfz"]), marcosize=1e6, seo=cases.oracle_seo())
    ref.run(-200e-9, DT, steps)
    err = cases.rel_err(B.LocalBunch, ref.M)
    print("main_track vs oracle", err, flush=True)
    assert max(err[0], err[4], err[9]) < 1e-8, err
    assert max(err.values()) < 1e-4, err
    assert np.array_equal(B.LocalBunch[:, 8] != 0, ref.M[:, 8] != 0)
    snaps = sorted(os.listdir(tmp_path / "out" / "snap"))
    assert len(snaps) == 3 and all(s.startswith("time_") for s in snaps)
    assert len(os.listdir(tmp_path / "out" / "tbt")) >= 1


def test_restart_from_a_step_dump_continues_the_run(tmp_path):
    from pyffa_amd import FFAG_Distribution as FD
    from pyffa_amd import main_track
    print("main_track restart", flush=True)
    steps = 4600
    snap = str(tmp_path / "snap")
    # one snapshot, when t first passes 15 ns: part of the painted bunch is injected by then
    dumps = [{"type": "StepDumpBunch", "start_time": 15e-9, "end_time": 17e-9, "interval_time": 10e-9,
              "save_folder": snap, "filtter_survive_paticles": False}]
    path, cfg = _setup(str(tmp_path), enable_sc=False, steps=steps, dumps=dumps)
    FD.set_seed(12)
    _, B_full, _ = main_track.main(path)
    files = os.listdir(snap)
    assert len(files) == 1
    dump = np.load(os.path.join(snap, files[0]))["particles"]
    t_dump = dump[0, -1]
    done = int(round((t_dump + 200e-9) / DT))
    inj = dump[:, 7] > 0.5
    assert 0 < inj.sum() < dump.shape[0]
    cfg["track"]["stop_condition"]["max_stepsN"] = steps - done
    cfg["DumpPara"]["modules"] = []
    path2 = str(tmp_path / "config_restart.json")
    with open(path2, "w") as f:
        json.dump(cfg, f)
    _, B_rest, _ = main_track.main(path2, restart_bunch=os.path.join(snap, files[0]))
    a = B_full.LocalBunch[np.argsort(B_full.LocalBunch[:, 15])]
    b = B_rest.LocalBunch[np.argsort(B_rest.LocalBunch[:, 15])]
    assert np.array_equal(a[:, 15], b[:, 15]) and np.array_equal(a[:, 7:9], b[:, 7:9])
    for c in (0, 1, 2, 3, 4, 5):
        np.testing.assert_allclose(b[:, c], a[:, c], rtol=1e-9, atol=1e-12 * np.abs(a[:, c]).max())
    np.testing.assert_allclose(np.mod(b[:, 9], 2 * np.pi), np.mod(a[:, 9], 2 * np.pi), atol=1e-9)


def test_sc_through_the_painted_injection_ramp_per_macro_step(tmp_path):
    """Strong space charge (1e9 protons per macro-particle) while the painted
    sub-bunches are injected: every 20-step macro step of the GPU run is
    compared with the oracle restarted from the GPU state (the empty solves
    before injection, the first few-particle solves and the newly injected
    particles' field history included)."""
    from pyffa_amd import FFAG_Distribution as FD
    from pyffa_amd import main_track
    print("sc through the injection ramp", flush=True)
    start, stop = 3700, 4500
    path, cfg = _setup(str(tmp_path), enable_sc=True, steps=start, dumps=[])
    FD.set_seed(13)
    rk, B, gp = main_track.main(path)
    t = rk.last_run["t"]
    tr = O.Tracker(cases.oracle_machine(gp), B.StaticBunchLocal.copy(), enable_sc=True, sc_step=20,
                   grid=tuple(cfg["SC"]["grid_size_rfz"]), marcosize=1e9, seo=cases.oracle_seo())
    worst = {}
    seen_bad = seen_good = False
    for k in range(start // 20, stop // 20):
        cases.sync_tracker_from_bunch(tr, B)
        tr.run(t, DT, (k + 1) * 20, start_step=k * 20)
        lp = dict(stop_condition=dict(max_stepsN=(k + 1) * 20, max_turn=1e9, max_time=1.0), time_step=DT,
                  step_dumps=None, enable_SC=True, SC_type=1, SC_step=20, verbose=False, start_step=k * 20)
        rk.rk4_solve_dt_bunch3_withSC_boris(None, t, B, gp, lp)
        t = rk.last_run["t"]
        seen_bad |= bool(tr.last_sc["bad_grid"])
        seen_good |= not tr.last_sc["bad_grid"]
        assert bool(tr.last_sc["bad_grid"]) == (B.sc_result()["bad_grid"] != 0.0)
        for c, v in cases.rel_err(B.LocalBunch, tr.M).items():
            worst[c] = max(worst.get(c, 0.0), v)
    print("per macro step worst", worst, flush=True)
    assert seen_bad and seen_good
    assert int((B.LocalBunch[:, 7] != 0).sum()) == 1200
    # r, phi, RF to FP64 round-off; the velocities carry the complex64 Poisson
    # solve's rounding (the oracle solves in FP64)
    assert max(worst[0], worst[4], worst[9]) < 1e-12, worst
    assert worst[5] < 1e-11 and worst[1] < 1e-9 and worst[2] < 1e-8 and worst[3] < 1e-7, worst
