"""Shared test inputs: one lattice, its Taylor maps, SEO tables, RF file and
seeded bunches, packaged both for the oracle (``oracle.Machine``) and for the
GPU path (``FFAG_GlobalParameters`` with the product's field objects).

Lattices: ``prescribed`` is the 12-sector scaling FFAG exactly as the
workload specifies it (D polarity -0.728), which is vertically unstable under
the reference field model, so it is only used for short (single-step /
few-step) parity checks; ``stable`` changes only the D polarity to -1.4
(tunes measured with the oracle: Qx, Qz fractional parts ~0.30, ~0.37;
vertical dynamic aperture ~2 cm) and is used for long runs."""
from __future__ import annotations

import functools
import os
import tempfile

import numpy as np

from oracle import oracle as O
from pyffa_amd import lattice as Lt

D_STABLE = -1.4


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
    """Seeded bunch converted to the tracker state layout."""
    from pyffa_amd.FFAG_ParasAndConversion import FFAG_ConversionTools
    P = Lt.gaussian_bunch(lattice(kind), seo(kind), n, seed=seed, inj_t_ns=inj_t_ns, fi_spread=fi_spread)
    ct = FFAG_ConversionTools()
    return ct.ConvertPrzek2Vrzek_boris(P) if solver == "boris" else ct.ConvertPrzek2Vrzek(P)


def bmap_config(lat):
    return lat.to_config()


def machines(kind="stable", max_order=2, rf=False, aperture=None, harmonics=None, tmpdir=None):
    """Returns (oracle Machine, product FFAG_GlobalParameters)."""
    from pyffa_amd.FFAG_Field import FFAG_Bfield_analytical, FFAG_BField_Error, FFAG_EField_new, FFAG_EField_spiral
    from pyffa_amd.FFAG_ParasAndConversion import FFAG_GlobalParameters
    lat = lattice(kind)
    bz, br, bfi = maps(kind, max(max_order, 1))
    bm = FFAG_Bfield_analytical.from_arrays(bmap_config(lat), max_order, (bz, br, bfi))
    gp = FFAG_GlobalParameters()
    gp.AddBMap(bm)
    mc = O.Machine(r_min=bm.r_min, r_step=bm.r_step, r_size=bm.r_size, fi_min=bm.fi_min, fi_step=bm.fi_step,
                   fi_size=bm.fi_size, Bz_stack=np.ascontiguousarray(bz[: max_order + 1]),
                   Br_stack=np.ascontiguousarray(br[: max_order + 1]),
                   Bfi_stack=np.ascontiguousarray(bfi[: max_order + 1]), max_order=max_order,
                   s_poly=np.ascontiguousarray(bm.s_rad_polynomial), sp_poly=np.ascontiguousarray(bm.s_rad_polynomial_prime),
                   B0=bm.B0, k_value=bm.k_value, nsectors=bm.Nsectors, convex=np.ascontiguousarray(bm.convex_params))
    if rf:
        tmpdir = tmpdir or tempfile.mkdtemp()
        path = emap_file(lat, tmpdir, **(rf if isinstance(rf, dict) else {}))
        spiral = isinstance(rf, dict) and rf.get("model") == "spiral"
        em = FFAG_EField_spiral(path, True) if spiral else FFAG_EField_new(path, True)
        gp.AddEMap(em)
        mc.emap_enable = True
        mc.gap_azimuths = em.gap_azimuths.copy()
        mc.gap_width = em.gap_width
        mc.rf_shift = em.rf_shift.copy()
        mc.freq_curve = em.freq_curve.copy()
        if spiral:
            mc.gap_model = "spiral"
            mc.emap_er, mc.emap_ez, mc.emap_ephi = (np.ascontiguousarray(a) for a in
                                                    (em.Er_map, em.Ezcoef_map, em.Ephi_map))
            mc.emap_r_axis, mc.emap_phi_axis = em.r_axis.copy(), em.phi_axis.copy()
        else:
            mc.e_rmin, mc.e_rmax = em.rmin, em.rmax
    if aperture is not None:
        gp.AddAperture(*aperture)
        mc.aperture = np.array(aperture, dtype=np.float64)
    if harmonics:
        bh = FFAG_BField_Error(harmonics, True)
        gp.AddBHarmonics(bh)
        mc.harm_enable = True
        mc.harm_m, mc.harm_amp, mc.harm_phase = bh.m_arr.copy(), bh.amp_arr.copy(), bh.phase_arr.copy()
        mc.harm_rmin, mc.harm_rmax = bh.rmin_arr.copy(), bh.rmax_arr.copy()
    return mc, gp


def oracle_machine(gp):
    """Oracle Machine holding the same fields as a product
    FFAG_GlobalParameters (maps, fixed or spiral gaps, aperture, harmonics)."""
    bm = gp.Bmap
    mo = bm.max_order
    mc = O.Machine(r_min=bm.r_min, r_step=bm.r_step, r_size=bm.r_size, fi_min=bm.fi_min, fi_step=bm.fi_step,
                   fi_size=bm.fi_size, Bz_stack=np.ascontiguousarray(bm.Bz_stack[: mo + 1]),
                   Br_stack=np.ascontiguousarray(bm.Br_stack[: mo + 1]),
                   Bfi_stack=np.ascontiguousarray(bm.Bfi_stack[: mo + 1]), max_order=mo,
                   s_poly=np.ascontiguousarray(bm.s_rad_polynomial), sp_poly=np.ascontiguousarray(bm.s_rad_polynomial_prime),
                   B0=bm.B0, k_value=bm.k_value, nsectors=bm.Nsectors, convex=np.ascontiguousarray(bm.convex_params))
    em = gp.Emap
    if em is not None and em.EnableFlag:
        mc.emap_enable = True
        mc.gap_azimuths = em.gap_azimuths.copy()
        mc.gap_width = em.gap_width
        mc.rf_shift = em.rf_shift.copy()
        mc.freq_curve = em.freq_curve.copy()
        if em.GapType == "spiral":
            mc.gap_model = "spiral"
            mc.emap_er, mc.emap_ez, mc.emap_ephi = (np.ascontiguousarray(a) for a in
                                                    (em.Er_map, em.Ezcoef_map, em.Ephi_map))
            mc.emap_r_axis, mc.emap_phi_axis = em.r_axis.copy(), em.phi_axis.copy()
        else:
            mc.e_rmin, mc.e_rmax = em.rmin, em.rmax
    if gp.Aperture_enable:
        mc.aperture = np.array(gp.Aperture_m, dtype=np.float64)
    bh = gp.BHarmonics
    if bh is not None and bh.enable_flag and bh.Nharm:
        mc.harm_enable = True
        mc.harm_m, mc.harm_amp, mc.harm_phase = bh.m_arr.copy(), bh.amp_arr.copy(), bh.phase_arr.copy()
        mc.harm_rmin, mc.harm_rmax = bh.rmin_arr.copy(), bh.rmax_arr.copy()
    return mc


def oracle_seo(kind="stable"):
    s = seo(kind)
    return O.SEO(s.Ek_axis, s.fi_axis, s.r_matrix, s.pr_matrix, s.perimeter)


APERTURE = (8.0, 9.3, -0.05, 0.05)     # the map's r range, |z| < 5 cm


def sync_tracker_from_bunch(tr, B):
    """Make an oracle Tracker continue exactly from a GPU bunch's state:
    state columns, flags, SC endpoint history and masks, injection head."""
    tr.M[:] = B.LocalBunch
    fl = B.ordered("flags").cpu().numpy()
    tr.E_pre[:] = B.ordered("esc_pre").cpu().numpy().astype(np.float64).T
    tr.E_prepre[:] = B.ordered("esc_prepre").cpu().numpy().astype(np.float64).T
    tr.m_pre[:] = (fl & 4) != 0
    tr.m_prepre[:] = (fl & 8) != 0
    tr.head = B._uninject_head
    tr.step_survive[:] = (fl & 2) != 0


def set_bunch_seo(B, kind="stable"):
    s = seo(kind)
    B.SEO_fi_axis, B.SEO_Ek_axis = s.fi_axis, s.Ek_axis
    B.SEO_r_matrix, B.SEO_pr_matrix, B.SEO_perimeter = s.r_matrix, s.pr_matrix, s.perimeter


STATE_COLS = (0, 1, 2, 3, 4, 5, 9)


def rel_err(a, b, cols=STATE_COLS):
    """Per-column max |a-b| / (|b| + rms(b)) over the state columns."""
    out = {}
    for c in cols:
        x, y = a[:, c], b[:, c]
        scale = np.abs(y) + np.sqrt(np.mean(y * y)) + 1e-300
        out[c] = float(np.max(np.abs(x - y) / scale))
    return out
