"""Tracking entry point: the examples/main_track.py flow on the GPU path.

``main(config_file, restart_bunch=None)`` reads the same config_track.json
schema (``BmapAndSEO``, ``Emap``, ``track``, ``SC``, ``BunchPara``,
``Paint``, ``Aperture``, ``DumpPara``, ``check``), builds the injected bunch
(``FFAG_Distribution.GenerateBunches``) or reads a restart file (an .npz
with ``particles`` or a CSV, rows in the dumps' (r, pr, z, pz, ...) layout),
loads the Bmap / SEO / Emap folders, configures the detectors and runs the
Boris (solver_type 1) or RK4 (solver_type 0) tracker
(examples/main_track.py:52-355).  As there, tracking starts at
t = -200 ns, or at the time stored in the restart file.

One optional key is added: ``track.precision`` = ``"fp64"`` (default, the
FP64 path) or ``"fp32"`` (the FP32-field tier, Boris only).

Command line: ``python -m pyffa_amd.main_track -j config_track.json
[-r restart.npz]``.
"""
from __future__ import annotations

import argparse
import json
import os

import numpy as np

from .FFAG_Bunch import FFAG_Bunch, FFAG_ManageBunchAttribute
from .FFAG_Distribution import GenerateBunches, LoadSEOParams
from .FFAG_Field import FFAG_Bfield_analytical, FFAG_BField_Error, FFAG_EField_new, FFAG_EField_spiral
from .FFAG_ParasAndConversion import FFAG_ConversionTools, FFAG_GlobalParameters
from .FFAG_dump import Dumps, PositionDump, PositionDumpBunch, StepDump, StepDumpBunch
from .FFAG_track import FFAG_RungeKutta, FunctionForAccelerationBunch_dt


def load_config_from_json(json_path):
    with open(json_path, "r") as f:
        return json.load(f)


def wrap_0_2pi(x):
    return x % (2.0 * np.pi)


def calc_injection_rf_phase(EMap, Azimuth_start, t_start, T0, Ek_inj, gap_idx=0):
    """RF phase at t_start that gives the design particle the phase acc_phi0
    when it reaches gap ``gap_idx`` (examples/main_track.py:24-47)."""
    phi_start = wrap_0_2pi(Azimuth_start)
    if EMap.GapType == "fixed":
        phi_gap = EMap.gap_azimuths[gap_idx]
    elif EMap.GapType == "spiral":
        phi_gap = EMap.phi_gap_interp_unwrapped(Ek_inj) + (EMap.gap_azimuths[gap_idx] - EMap.gap_azimuths[0])
    else:
        raise ValueError("EMap.type must be 'fixed' or 'spiral'")
    dphi_orbit = phi_gap - phi_start
    return EMap.acc_phi0 - EMap.harmonic * dphi_orbit - EMap.harmonic * 2.0 * np.pi * (0.0 - t_start) / T0


def build_dumps(config):
    """Detectors from ``DumpPara.modules`` (one per azimuth for the position
    detectors)."""
    dumps = Dumps()
    for d in config["DumpPara"]["modules"]:
        kind = d["type"]
        if kind == "StepDump":
            dumps.add_dump(StepDump(d["start_time"], d["end_time"], d["interval_time"],
                                    d["tracked_particle_id_global"], save_folder=d["save_folder"]))
        elif kind == "PositionDump":
            for az in d["dump_azimuth"]:
                dumps.add_dump(PositionDump(az, d["start_time"], d["end_time"], d["num_particles_to_dump_global"],
                                            save_folder=d["save_folder"]))
        elif kind == "StepDumpBunch":
            dumps.add_dump(StepDumpBunch(d["start_time"], d["end_time"], d["interval_time"],
                                         save_folder=d["save_folder"],
                                         filter_survive_flag=d["filtter_survive_paticles"]))
        elif kind == "PositionDumpBunch":
            for az in d["dump_azimuth"]:
                dumps.add_dump(PositionDumpBunch(az, d["start_time"], d["end_time"], save_folder=d["save_folder"]))
    return dumps


def main(config_file, restart_bunch=None):
    """Runs one tracking job; returns (solver, bunch, global parameters) for
    callers that want to inspect the final state."""
    config = load_config_from_json(config_file)
    bmap_dir = config["BmapAndSEO"]["maps"]
    max_order = config["BmapAndSEO"]["max_order"]
    seo_dir = os.path.join(bmap_dir, config["BmapAndSEO"]["SEO_FileName"])
    emap_path = config["Emap"]["maps"]
    ek_mev = config["track"]["start_EkMeV"]
    azimuth_start = config["track"]["start_azimuth"]
    precision = config["track"].get("precision", "fp64")
    check = config["check"]
    paint = config["Paint"]
    bp = dict(config["BunchPara"])
    bp.update(PlotFlag=check["CheckBunch"], PlotRMS=check["PlotRMS"], BmapPATHName=bmap_dir, SEO=seo_dir,
              InjectEk=ek_mev, InjPosition=azimuth_start, PaintEnable=paint["enable"],
              PaintMaxNum=paint["MaxBunchNum"], PaintTimeInterval=paint["TimeInterval"], PaintCurve=paint["Curve"],
              EmapPATHName=emap_path)
    conv = FFAG_ConversionTools()
    attr = FFAG_ManageBunchAttribute().Attribute
    if restart_bunch is not None:
        if not os.path.exists(restart_bunch):
            raise FileNotFoundError(f"Error: restart_bunch file '{restart_bunch}' not found.")
        if restart_bunch.endswith(".npz"):
            dist_p = np.load(restart_bunch)["particles"]
        else:
            dist_p = np.loadtxt(restart_bunch, delimiter=",")
        dist_p = dist_p[np.argsort(dist_p[:, attr["inj_t"]], kind="stable")]
        dist_v = conv.ConvertPrzek2Vrzek(dist_p)
        dist_v_boris = conv.ConvertPrzek2Vrzek_boris(dist_p)
    else:
        dist_p = GenerateBunches(bp)
        dist_p = dist_p[np.argsort(dist_p[:, attr["inj_t"]], kind="stable")]
        dist_v = conv.ConvertPrzek2Vrzek(dist_p)
        dist_v_boris = conv.ConvertPrzek2Vrzek_boris(dist_p)
        gid = np.arange(dist_p.shape[0])
        dist_v[:, attr["Global_ID"]] = gid
        dist_v_boris[:, attr["Global_ID"]] = gid

    dumps = build_dumps(config)
    seo_inj, bmap_path, seo_info = LoadSEOParams(os.path.join(seo_dir, "SEO_ini.txt"), ek_mev)
    seo_r = np.loadtxt(os.path.join(seo_dir, "SEO_r.txt"), skiprows=1)
    seo_pr = np.loadtxt(os.path.join(seo_dir, "SEO_pr.txt"), skiprows=1)
    gp = FFAG_GlobalParameters()
    bmap = FFAG_Bfield_analytical(os.path.join(bmap_path, "config_Bmap.json"), max_order)
    header, _ = FFAG_EField_new.ParseEmapFile(emap_path)
    gap_type = str(header["gap_model"]).strip().lower()
    if gap_type == "fixed":
        emap = FFAG_EField_new(emap_path, config["Emap"]["enable"])
    elif gap_type == "spiral":
        emap = FFAG_EField_spiral(emap_path, config["Emap"]["enable"])
    else:
        raise ValueError(f"Unknown Emap gap type: {gap_type!r}")
    with open(os.path.join(seo_dir, "config_SEO.json")) as f:
        seo_cfg = json.load(f)
    gp.AddBMap(bmap)
    gp.AddEMap(emap)
    gp.AddSEOInfo(seo_info)
    gp.AddBHarmonics(FFAG_BField_Error(seo_cfg["BHarmonics"]["harmonics"], seo_cfg["BHarmonics"]["enable"]))
    ap = config["Aperture"]
    if ap["enable"]:
        gp.AddAperture(ap["Rmin"], ap["Rmax"], ap["Zmin"], ap["Zmax"])

    t_start = -200.0e-9
    lp = {"stop_condition": config["track"]["stop_condition"], "time_step": config["track"]["time_step"],
          "step_dumps": dumps, "enable_SC": config["SC"]["enable_SC"], "SC_type": config["SC"]["SC_type"],
          "SC_step": config["SC"]["step"], "precision": precision}
    grid = tuple(config["SC"]["grid_size_rfz"])
    solver_type = config["track"]["solver_type"]
    if restart_bunch is not None:
        t_start = float(dist_p[0, -1])
    else:
        inj_phase = calc_injection_rf_phase(emap, azimuth_start, t_start, seo_inj[3], seo_inj[0], gap_idx=0)
        dist_v[:, 9] = inj_phase
        dist_v_boris[:, 9] = inj_phase

    rk = FFAG_RungeKutta()
    if solver_type not in (0, 1):
        raise ValueError("solver_type must be 0 or 1")
    bunch = FFAG_Bunch(dist_v if solver_type == 0 else dist_v_boris, marcosize=bp["ParticleDensity"],
                       sc_grid_size=grid, BunchType="RK4" if solver_type == 0 else "Boris", precision=precision)
    bunch.SEO_fi_axis = seo_r[0, 1:]
    bunch.SEO_Ek_axis = seo_r[1:, 0]
    bunch.SEO_r_matrix = seo_r[1:, 1:]
    bunch.SEO_pr_matrix = seo_pr[1:, 1:]
    bunch.SEO_perimeter = seo_info[:, 8]
    if restart_bunch is not None:
        bunch.restore_from_restart_matrix()
    if solver_type == 0:
        rk.rk4_solve_dt_bunch3_withSC(FunctionForAccelerationBunch_dt, t_start, bunch, gp, lp)
    else:
        rk.rk4_solve_dt_bunch3_withSC_boris(FunctionForAccelerationBunch_dt, t_start, bunch, gp, lp)
    return rk, bunch, gp


if __name__ == "__main__":
    parser = argparse.ArgumentParser(description="FFAG multiparticle tracking on the GPU")
    parser.add_argument("-j", type=str, default="./config_track.json", help="config_track.json path")
    parser.add_argument("-r", type=str, default=None, help="restart bunch file (optional)")
    args = parser.parse_args()
    main(args.j, args.r)
