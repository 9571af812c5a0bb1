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
