"""Synthetic workloads of the tracking path: one lattice, its Taylor maps,
closed-orbit (SEO) tables, RF file and seeded bunches, packaged as the
product's ``FFAG_GlobalParameters`` and ``FFAG_Bunch`` inputs.

The reference ships no maps, SEO tables or RF files, so every workload is
generated here (the same files ``main_Bmap`` / ``main_Emap`` / the SEO search
would write).  Lattices: ``prescribed`` is the 12-sector scaling FFAG with
the workload's D polarity -0.728, vertically unstable under the reference
field model (short runs only); ``stable`` changes only the D polarity to -1.4
(fractional tunes ~0.30 / ~0.37, vertical dynamic aperture ~2 cm)."""
from __future__ import annotations

import functools
import os
import tempfile

import numpy as np

from . import lattice as Lt

D_STABLE = -1.4
APERTURE = (8.0, 9.3, -0.05, 0.05)     # the map's r range, |z| < 5 cm


@functools.lru_cache(maxsize=None)
def lattice(kind="stable", rstep=0.001, theta_step_deg=0.005):
    pol_d = -0.728 if kind == "prescribed" else D_STABLE
    lat = Lt.ScalingFFAG(polarities=(0.0, 1.0, 0.0, pol_d, 0.0), rstep=rstep, theta_step_deg=theta_step_deg)
    Lt.solve_B0(lat, iters=2)
    return lat


@functools.lru_cache(maxsize=None)
def maps(kind="stable", max_order=2, rstep=0.001, theta_step_deg=0.005):
    return Lt.taylor_maps(lattice(kind, rstep, theta_step_deg), max_order)


@functools.lru_cache(maxsize=None)
def seo(kind="stable"):
    return Lt.seo_tables(lattice(kind), np.arange(285.0, 316.0, 5.0), n_theta=1000)


def emap_file(lat, tmpdir, v0=100e3, gap_width=0.2, ramp=(1.0, 1.002), model="fixed"):
    """4-column RF frequency-curve file (fixed rectangles or spiral maps):
    two gaps 180 degrees apart, harmonic 1, a linear frequency ramp."""
    co = Lt.closed_orbit(lat, lat.ek_ref_mev, n_theta=1000)
    f0 = 1.0 / co["T0"]
    if model == "spiral":
        path = os.path.join(tmpdir, "spiral", "FrequencyCurve_spiral.txt")
        Lt.write_spiral_emap(path, gap_azimuth_deg=(15.0, 195.0), harmonic=1, v0_volt=v0, f_start=f0 * ramp[0],
                             f_end=f0 * ramp[1], t_start=-300e-9, t_end=2e-3, n_t=2001, gap_width=0.05,
                             rmin=lat.rmin, rmax=lat.rmax)
        return path
    path = os.path.join(tmpdir, "FrequencyCurve_2gaps.txt")
    Lt.write_emap(path, gap_azimuth_deg=(0.0, 180.0), harmonic=1, v0_volt=v0, f_start=f0 * ramp[0],
                  f_end=f0 * ramp[1], t_start=-300e-9, t_end=2e-3, n_t=2001, gap_width=gap_width,
                  e_rmin=lat.rmin, e_rmax=lat.rmax)
    return path


def bunch(kind="stable", n=1000, seed=0, solver="boris", inj_t_ns=0.0, fi_spread=2 * np.pi):
    """Seeded Gaussian bunch about the closed orbit, converted to the
    tracker's (N, 16) state layout."""
    from .FFAG_ParasAndConversion import FFAG_ConversionTools
    P = Lt.gaussian_bunch(lattice(kind), seo(kind), n, seed=seed, inj_t_ns=inj_t_ns, fi_spread=fi_spread)
    ct = FFAG_ConversionTools()
    return ct.ConvertPrzek2Vrzek_boris(P) if solver == "boris" else ct.ConvertPrzek2Vrzek(P)


def global_parameters(kind="stable", max_order=2, rf=False, aperture=None, harmonics=None, tmpdir=None):
    """``FFAG_GlobalParameters`` of the lattice: order-0 analytic field plus
    Taylor maps up to ``max_order``, optional RF gaps (``rf``: False, True or
    a dict of ``emap_file`` options), aperture and harmonic errors."""
    from .FFAG_Field import FFAG_Bfield_analytical, FFAG_BField_Error, FFAG_EField_new, FFAG_EField_spiral
    from .FFAG_ParasAndConversion import FFAG_GlobalParameters
    lat = lattice(kind)
    bz, br, bfi = maps(kind, max(max_order, 1))
    bm = FFAG_Bfield_analytical.from_arrays(lat.to_config(), max_order, (bz, br, bfi))
    gp = FFAG_GlobalParameters()
    gp.AddBMap(bm)
    if rf:
        tmpdir = tmpdir or tempfile.mkdtemp()
        path = emap_file(lat, tmpdir, **(rf if isinstance(rf, dict) else {}))
        spiral = isinstance(rf, dict) and rf.get("model") == "spiral"
        gp.AddEMap(FFAG_EField_spiral(path, True) if spiral else FFAG_EField_new(path, True))
    if aperture is not None:
        gp.AddAperture(*aperture)
    if harmonics:
        gp.AddBHarmonics(FFAG_BField_Error(harmonics, True))
    return gp


def set_bunch_seo(B, kind="stable"):
    """Attach the lattice's SEO tables to a bunch (what main_track does from
    the SEO files)."""
    s = seo(kind)
    B.SEO_fi_axis, B.SEO_Ek_axis = s.fi_axis, s.Ek_axis
    B.SEO_r_matrix, B.SEO_pr_matrix, B.SEO_perimeter = s.r_matrix, s.pr_matrix, s.perimeter


# ---------------------------------------------------------------------------
# Spiral-sector acceleration cycle: a complete config_track.json job
# ---------------------------------------------------------------------------
SPIRAL_DEG = 35.0


def spiral_cycle_job(folder, *, n_particles, grid=(1024, 512, 512, 4.0), max_steps=5000, precision="fp32",
                     max_order=6, v0=200e3, ek_ext=600.0, turns=5000, paint_bunches=100, paint_interval_ns=None,
                     paint_offsets=(0.01, 0.0, 0.0, -1e-3), bunch_length_ns=None, protons=1e13,
                     seo_energies=None, enable_sc=True, aperture=True, emittance=(10.0, 1.0),
                     start_azimuth_deg=0.0):
    """Writes a complete job for main_track.main into ``folder``: the 12-sector
    scaling FFAG with a 35 degree spiral and the workload's polarities
    (0, 1, 0, -0.728, 0) (the spiral's edge focusing makes the -0.728 D
    vertically stable: Qr ~ 3.38, Qz ~ 0.45 at all energies), and

    * ``Bmap/``: order-``max_order`` Taylor maps from the GPU generator
      (``main_Bmap``) on the r 8.0-9.3 m (1 mm) x one-sector (0.005 deg) grid;
    * ``Bmap/resultsSEO/``: closed orbits, tunes and SEO tables of
      ``seo_energies`` (GPU closed-orbit search), Twiss functions;
    * ``Emap/``: two spiral RF gaps (``main_Emap``: gaps along the 35 degree
      spiral, the frequency curve of the synchronous particle at 30 degrees
      from 300 MeV to ``ek_ext`` over at most ``turns`` turns, ``v0`` per gap;
      the Gaussian gap profile integrates to U = 1 across the line and
      gap_width = 1 m, so a crossing gains V0, as in the reference's
      examples/config_Emap.py);
    * the painting curve: ``paint_bunches`` sub-bunches every
      ``paint_interval_ns`` (default one revolution, i.e. painting over the
      first ``paint_bunches`` turns) whose (x, px, y, py) offsets grow linearly
      to ``paint_offsets`` -- r up, pz down;
    * ``config_track.json``: ``n_particles`` macro-particles in total for
      ``protons`` protons, Gaussian (r, z) rms emittances ``emittance``
      [pi mm mrad], the SC grid, Boris at 0.05 ns, ``max_steps`` steps.  The
      default vertical emittance keeps the bunch (beta_z ~ 20 m) inside the
      ~2 cm where the Taylor z expansion converges at the D-magnet fringe.
      The injection point is at ``start_azimuth_deg``: the bunch generator
      takes the Twiss functions at the injection azimuth in radians read as
      degrees (src/FFAG_Distribution.py:1441), so only an injection at 0
      degrees gets the Twiss functions of its own azimuth, which matters in
      this lattice (alpha_z is -0.9 at 0 degrees and +3.2 at 15).

    Returns (config path, config dict)."""
    import json
    from . import FFAG_Emap as FE
    from . import FFAG_HighOrder as H
    from .FFAG_Field import FFAG_Bfield_analytical
    from .FFAG_ParasAndConversion import FFAG_GlobalParameters
    from .FFAG_SEO import FFAG_BetaFuncCalc, FFAG_SearchSEO
    os.makedirs(folder, exist_ok=True)
    lat = Lt.ScalingFFAG(polarities=(0.0, 1.0, 0.0, -0.728, 0.0), spiral_deg=SPIRAL_DEG)
    Lt.solve_B0(lat, iters=2)
    cfg_b = lat.to_config()
    cfg_b["config"] = {"folder": "Bmap", "expand_order": int(max_order), "date": 0.0}
    path_b = os.path.join(folder, "config_Bmap.json")
    with open(path_b, "w") as f:
        json.dump(cfg_b, f)
    bmap_dir = H.main_Bmap(path_b)
    bm = FFAG_Bfield_analytical(os.path.join(bmap_dir, "config_Bmap.json"), 0)
    gp = FFAG_GlobalParameters()
    gp.AddBMap(bm)
    gp.SEO_SaveFold = "resultsSEO"
    ek_list = np.arange(280.0, ek_ext + 41.0, 20.0) if seo_energies is None else np.asarray(seo_energies)
    seo_dir = FFAG_SearchSEO(gp).SearchSEOsControllerVect(ek_list)
    FFAG_BetaFuncCalc().CalcBetaFunc(seo_dir, ek_list, gp)
    sd = np.loadtxt(os.path.join(seo_dir, "SEO_ini.txt"), skiprows=2, ndmin=2)
    t_rev_ns = 1e9 / float(np.interp(300.0, sd[:, 1], sd[:, 4]))
    emap_dir = os.path.join(folder, "Emap")
    emap_file = os.path.join(emap_dir, "FrequencyCurve_spiral.txt")
    cfg_e = {
        "config": {"build_frequency_curve": True, "build_spiral_map": True, "filename": emap_file},
        "machine": {"energy_inj": 300.0, "energy_ext": float(ek_ext)},
        "Emap": dict(Bmap_fold=bmap_dir, SEO_FileName="resultsSEO", gap_model="spiral", gap_azimuth=[0.0],
                     gap_width=1.0, rmin=8.0, rmax=9.3, harmonic=1, Ngap=2, t_start=-300e-9,
                     spiral_rmin_calc=8.05, spiral_rmax_calc=9.25, acc_phase_start=30.0, acc_phase_end=30.0,
                     phase_ramp_start=0, phase_ramp_end=1, num_turns=int(turns), acc_voltage_start=float(v0),
                     acc_voltage_end=float(v0), voltage_ramp_start=0, voltage_ramp_end=1,
                     spiral_alpha_deg=SPIRAL_DEG, spiral_r_ref=8.5, spiral_s_ref_deg=20.0, spiral_Nphi=4000,
                     spiral_U=1.0, spiral_FWHM=0.05, spiral_rmin_save=8.05, spiral_rmax_save=9.25, spiral_nr=241,
                     spiral_nphi_full=7200, spiral_pad_deg=3.0, spiral_min_nphi=200, spiral_newton_iter=10,
                     spiral_normal_side=1, emap_save_dir=emap_dir, Er_map_file="Er_coef.txt",
                     Ez_map_file="Ez_coef.txt", Ephi_map_file="Ephi_coef.txt"),
    }
    path_e = os.path.join(folder, "config_Emap.json")
    with open(path_e, "w") as f:
        json.dump(cfg_e, f)
    FE.main_Emap(path_e)
    interval = float(paint_interval_ns) if paint_interval_ns is not None else t_rev_ns
    nb = max(1, int(paint_bunches))
    curve = os.path.join(folder, "paint_curve.txt")
    tc = np.array([0.0, nb * interval])
    off = np.array(paint_offsets, dtype=np.float64)
    np.savetxt(curve, np.column_stack([tc, np.outer([0.0, 1.0], off)]), header="t x px y py")
    t_len = float(bunch_length_ns) if bunch_length_ns is not None else 0.8 * interval
    cfg = {
        "config": {"date": 0.0, "MPI_process_num": 1, "thread_num": 1},
        "track": {"start_EkMeV": 300.0, "time_start": -200e-9, "time_step": 0.05e-9,
                  "start_azimuth": float(start_azimuth_deg),
                  "stop_condition": {"max_stepsN": int(max_steps), "max_turn": 1e9, "max_time": 1.0},
                  "solver_type": 1, "precision": precision},
        "SC": {"enable_SC": bool(enable_sc), "SC_type": 1, "grid_size_rfz": list(grid), "step": 20},
        "BmapAndSEO": {"maps": bmap_dir, "max_order": int(max_order), "SEO_FileName": "resultsSEO"},
        "Emap": {"maps": emap_file, "enable": True},
        "BunchPara": {"ParticleDensity": protons / n_particles, "ParticleNum": int(n_particles) // nb,
                      "TransverseREmit": float(emittance[0]), "TransverseZEmit": float(emittance[1]),
                      "LongitudeT": t_len, "LongitudeDEk": 1.0,
                      "InjTimeNanoSec": 0, "TransverseDistType": "gauss", "LongitudeDistType": "UniformGauss"},
        "Paint": {"enable": nb > 1, "MaxBunchNum": nb, "TimeInterval": interval, "Curve": curve},
        "Aperture": {"enable": bool(aperture), "Rmin": APERTURE[0], "Rmax": APERTURE[1], "Zmin": APERTURE[2],
                     "Zmax": APERTURE[3]},
        "DumpPara": {"modules": []},
        "check": {"CheckBunch": False, "PlotRMS": [4, 4, 4]},
    }
    path = os.path.join(folder, "config_track.json")
    with open(path, "w") as f:
        json.dump(cfg, f, indent=1)
    return path, cfg
