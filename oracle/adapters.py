"""Adapters from the product's host objects to the oracle's inputs -- TEST
INFRASTRUCTURE ONLY (tests, smoke() and bench.py's CPU-baseline leg).

``machine_from(gp)`` builds the oracle ``Machine`` holding the same fields
as a product ``FFAG_GlobalParameters`` (Taylor maps, fixed or spiral gaps,
aperture, harmonics); ``seo_from(seo)`` the oracle ``SEO`` of a product SEO
table set.  Nothing here imports the product package: the objects are
passed in.
"""
from __future__ import annotations

import numpy as np

from . import oracle as O


def machine_from(gp):
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


def seo_from(s):
    """Oracle SEO tables from a product SEO table set (Ek and azimuth axes,
    r0 and pr0 matrices, perimeter)."""
    return O.SEO(s.Ek_axis, s.fi_axis, s.r_matrix, s.pr_matrix, s.perimeter)
