"""CPU tests of the Emap builder: the spiral-gap map oracle against the
geometry of the log spiral (distance zero on the line, exactly d at d along
the normal, the profile integrating to U, the normal's side), and the
frequency-curve builder against the closed forms of fixed and spiral gaps."""
import json
import os

import numpy as np
import pytest

from oracle import emapgen as EG

R_REF, S_REF, ALPHA = 8.5, np.deg2rad(20.0), np.deg2rad(50.0)
A = np.tan(ALPHA)
RMIN, RMAX = 7.5, 14.0


def _line(r):
    s = S_REF + A * np.log(r / R_REF)
    cs, sn = np.cos(s), np.sin(s)
    t = np.stack([cs - A * sn, sn + A * cs]) / np.sqrt(1 + A * A)
    return np.stack([r * cs, r * sn]), t


def test_nodes_on_and_off_the_line():
    r = np.linspace(8.0, 13.0, 41)
    p, t = _line(r)
    n = np.stack([-t[1], t[0]])                                   # the left normal
    for d in (0.0, 0.05, -0.12, 0.3):
        q = p + d * n
        R, PHI = np.hypot(q[0], q[1]), np.unwrap(np.arctan2(q[1], q[0]))
        dist, ux, uy, flag = EG.nearest(R, PHI, R_REF, S_REF, A, RMIN, RMAX)
        np.testing.assert_allclose(dist, abs(d), atol=1e-10)
        assert np.all(flag == 0)
        np.testing.assert_allclose(ux, n[0], atol=1e-10)
        np.testing.assert_allclose(uy, n[1], atol=1e-10)


def test_field_on_the_line_and_profile_integral():
    U, fwhm = 1.7, 0.6
    delta = fwhm / 2.35482
    r = np.array([9.0, 11.0])
    p, t = _line(r)
    n = np.stack([-t[1], t[0]])
    phi = S_REF + A * np.log(r / R_REF)
    er, ef, _, d, _ = EG.maps(r, phi, R_REF, S_REF, A, RMIN, RMAX, U, delta)
    peak = U / (delta * np.sqrt(2 * np.pi))
    for i in range(2):
        c, s = np.cos(phi[i]), np.sin(phi[i])
        assert abs(d[i, i]) < 1e-12
        np.testing.assert_allclose([er[i, i], ef[i, i]], peak * np.array([n[0, i] * c + n[1, i] * s,
                                                                          -n[0, i] * s + n[1, i] * c]), rtol=1e-12)
    # across the line: the field magnitude integrates to U
    ds = np.linspace(-6 * delta, 6 * delta, 4001)
    q = p[:, :1] + ds[None, :] * n[:, :1]
    R, PHI = np.hypot(q[0], q[1]), np.unwrap(np.arctan2(q[1], q[0]))
    dist, ux, uy, flag = EG.nearest(R, PHI, R_REF, S_REF, A, RMIN, RMAX)
    eh = U / (delta * np.sqrt(2 * np.pi)) * np.exp(-0.5 * (dist / delta) ** 2)
    assert abs(np.trapezoid(eh, ds) / U - 1) < 1e-6


def test_planted_normal_side_slip_is_caught():
    """side = -1 gives the right normal: the orientation check of the first
    test (left normal) fails for it."""
    r = np.linspace(8.0, 13.0, 5)
    p, t = _line(r)
    phi = S_REF + A * np.log(r / R_REF)
    _, ux, uy, _ = EG.nearest(r, phi, R_REF, S_REF, A, RMIN, RMAX, side=-1.0)
    assert np.all(ux * (-t[1]) + uy * t[0] < -0.99)


def test_end_flags():
    r = np.array([7.0, 15.0])
    phi = S_REF + A * np.log(r / R_REF)
    _, _, _, flag = EG.nearest(r, phi, R_REF, S_REF, A, RMIN, RMAX)
    assert flag.tolist() == [1, 2]


def _seo_folder(tmp_path, ek, f_rev, r_of_ek):
    d = tmp_path / "Bmap" / "resultsSEO_noError"
    os.makedirs(d)
    data = np.zeros((ek.size, 17))
    data[:, 0], data[:, 1], data[:, 4], data[:, 6] = np.arange(ek.size), ek, f_rev, r_of_ek
    np.savetxt(d / "SEO_ini.txt", data, header="Bmap foldname=x\ntitles")
    fi = np.linspace(0, 2 * np.pi, 721)
    m = np.zeros((ek.size + 1, fi.size + 1))
    m[0, 1:], m[1:, 0], m[1:, 1:] = fi, ek, r_of_ek[:, None]
    np.savetxt(d / "SEO_r.txt", m, header="unitEk=MeV, unitFi=deg, unitR=m")
    return str(tmp_path / "Bmap")


def _emap_cfg(bmap, model, **kw):
    e = dict(Bmap_fold=bmap, gap_model=model, gap_azimuth=[0.0, 180.0], gap_width=0.2, rmin=8.0, rmax=9.3,
             harmonic=1, Ngap=2, t_start=-200e-9, spiral_rmin_calc=RMIN, spiral_rmax_calc=RMAX,
             acc_phase_start=30.0, acc_phase_end=30.0, phase_ramp_start=1000, phase_ramp_end=2000, num_turns=400,
             acc_voltage_start=100e3, acc_voltage_end=100e3, voltage_ramp_start=1000, voltage_ramp_end=1100,
             spiral_alpha_deg=50.0, spiral_r_ref=R_REF, spiral_s_ref_deg=20.0, spiral_Nphi=4000)
    e.update(kw)
    return e


def test_frequency_curve_fixed_gaps_closed_form(tmp_path):
    from pyffa_amd import FFAG_Emap as FE
    from pyffa_amd.FFAG_Field import FFAG_EField_new
    ek = np.array([280.0, 300.0, 320.0, 340.0])
    f_rev = 3.5e6 + 2e3 * (ek - 280.0)                            # linear in Ek
    bmap = _seo_folder(tmp_path, ek, f_rev, np.full(4, 8.5))
    path = str(tmp_path / "freq.txt")
    out = FE.build_frequency_curve_from_config({"filename": path}, _emap_cfg(bmap, "fixed", num_turns=100),
                                               {"energy_inj": 300.0, "energy_ext": 600.0})
    t, f, v, e = out.T
    assert np.all(np.diff(t) > 0) and np.allclose(np.diff(t), np.diff(t)[0])
    assert abs((t[1] - t[0]) - 0.5 / f.max()) < 1e-20
    # 200 gaps of V0 sin 30 deg: 0.05 MeV each; at constant phase each event's
    # RF frequency is h f_rev of the energy after its gap, so f = f_rev(E)
    # along the curve (both resampled linearly in t)
    assert abs(e[-1] - (300.0 + 200 * 0.05)) < 1e-9
    np.testing.assert_allclose(f, 3.5e6 + 2e3 * (e - 280.0), rtol=1e-12)
    assert np.all(v == 100e3)
    em = FFAG_EField_new(path, True)
    assert np.allclose(em.gap_azimuths, np.deg2rad([0.0, 180.0])) and np.allclose(em.rf_shift, np.deg2rad([0, -180.0]))
    assert em.harmonic == 1 and em.freq_curve.shape[1] == 4


def test_frequency_curve_phase_ramp_term(tmp_path):
    from pyffa_amd import FFAG_Emap as FE
    ek = np.array([280.0, 300.0, 320.0])
    bmap = _seo_folder(tmp_path, ek, np.full(3, 3.6e6), np.full(3, 8.5))
    out = FE.build_frequency_curve_from_config(
        {"filename": str(tmp_path / "f.txt")},
        _emap_cfg(bmap, "fixed", num_turns=30, phase_ramp_start=10, phase_ramp_end=20, acc_phase_end=40.0),
        {"energy_inj": 300.0, "energy_ext": 600.0})
    t, f = out[:, 0], out[:, 1]
    # f_rev constant: the RF frequency is h f_rev + dphi / (2 pi dt) with
    # dphi = 10 deg / 20 gap events during the ramp
    dt = 1 / 3.6e6 / 2
    ramp = 3.6e6 + np.deg2rad(10.0) / 20 / (2 * np.pi * dt)
    assert abs(f.max() - ramp) / ramp < 1e-9 and abs(f[-1] - 3.6e6) < 1e-3


def test_spiral_gap_azimuths_from_circular_orbits(tmp_path):
    from pyffa_amd import FFAG_Emap as FE
    ek = np.array([280.0, 300.0, 320.0, 340.0])
    radii = np.array([8.6, 9.2, 10.1, 11.3])
    bmap = _seo_folder(tmp_path, ek, np.full(4, 3.6e6), radii)
    f, ek_n, phi_n, r_n = FE.build_phi_gap_interpolator_from_SEO(
        os.path.join(bmap, "resultsSEO_noError", "SEO_r.txt"), 50.0, R_REF, 20.0, RMIN, RMAX, 4000)
    np.testing.assert_allclose(r_n, radii, rtol=1e-9)
    np.testing.assert_allclose(phi_n, S_REF + A * np.log(radii / R_REF), atol=1e-9)
    assert abs(f(310.0) - (phi_n[1] + phi_n[2]) / 2) < 1e-12
    out = FE.build_frequency_curve_from_config({"filename": str(tmp_path / "s.txt")},
                                               _emap_cfg(bmap, "spiral", num_turns=50),
                                               {"energy_inj": 300.0, "energy_ext": 600.0})
    assert out[-1, 3] > 300.0
    with open(tmp_path / "s.txt") as fh:
        head = fh.read(4000)
    assert "gap_model = spiral" in head and "phi_gap_unwrapped_deg" in head
