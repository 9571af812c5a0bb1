"""CPU tests of the spiral RF gap model (FFAG_EField_spiral): the oracle's
window search and bilinear interpolation pinned by closed forms, and the
file format of the spiral Emap files read back by the product's class."""
import numpy as np

from oracle import oracle as O

TWO_PI = 2 * np.pi


def _affine(a, b, c, d):
    return lambda r, p: a + b * r + c * p + d * r * p


def _eval(M, gaz, rf_amp, r_axis, p_axis, er, ez, ef):
    n = M.shape[0]
    Er, Ez, Efi = np.zeros(n), np.zeros(n), np.zeros(n)
    gap = np.zeros(n, np.int32)
    gaz = np.ascontiguousarray(gaz, dtype=np.float64)
    O.lib().ora_emap_spiral(O._p(M), O._i64(n), O._p(gaz), O._i(gaz.size), O._d(rf_amp), O._d(r_axis[0]),
                            O._d(r_axis[1] - r_axis[0]), O._i64(r_axis.size), O._d(p_axis[0]),
                            O._d(p_axis[1] - p_axis[0]), O._i64(p_axis.size), O._p(er), O._p(ez), O._p(ef),
                            O._p(Er), O._p(Ez), O._p(Efi), O._p(gap))
    return Er, Ez, Efi, gap


def _grid():
    r_axis = np.linspace(8.0, 9.0, 21)
    p_axis = np.linspace(-0.2, 0.3, 26)
    R, P = np.meshgrid(r_axis, p_axis, indexing="ij")
    fr, fz, ff = _affine(0.3, -1.5, 2.0, 0.7), _affine(-2.0, 0.5, 1.0, -0.3), _affine(1.0, 0.2, -3.0, 0.9)
    maps = [np.ascontiguousarray(f(R, P)) for f in (fr, fz, ff)]
    return r_axis, p_axis, (fr, fz, ff), maps


def _particles(r, fi, z):
    M = np.zeros((r.size, 16))
    M[:, 0], M[:, 2], M[:, 4] = r, z, fi
    M[:, 7] = M[:, 8] = 1.0
    return M


def test_bilinear_is_exact_on_bilinear_maps_inside_and_outside():
    """A map a + b r + c phi + d r phi is reproduced exactly (bilinear weights,
    also where clamped indices extrapolate beyond the map)."""
    r_axis, p_axis, fs, maps = _grid()
    rng = np.random.default_rng(0)
    n = 5000
    r = rng.uniform(7.8, 9.2, n)                       # partly outside the r range
    loc = rng.uniform(-0.2, 0.3, n)
    z = rng.uniform(-0.02, 0.02, n)
    gaz = np.array([1.0])
    turns = rng.integers(-3, 4, n)
    fi = gaz[0] + loc + TWO_PI * turns
    M = _particles(r, fi, z)
    amp = 1234.5
    Er, Ez, Efi, gap = _eval(M, gaz, amp, r_axis, p_axis, *maps)
    # the window search moves the azimuth back by whole turns (rounding ~1e-15)
    np.testing.assert_allclose(Er, amp * fs[0](r, loc), rtol=1e-11, atol=1e-9)
    np.testing.assert_allclose(Ez, amp * fs[1](r, loc) * z, rtol=1e-11, atol=1e-9)
    np.testing.assert_allclose(Efi, amp * fs[2](r, loc), rtol=1e-11, atol=1e-9)
    assert np.all(gap == 0)


def test_window_search_first_hit_and_misses():
    r_axis, p_axis, fs, maps = _grid()
    gaz = np.array([0.5, 0.6, 3.0])                    # windows of gaps 0 and 1 overlap
    r = np.full(6, 8.5)
    fi = np.array([0.5 + 0.25, 0.6 + 0.29, 3.0 + 0.1, 3.0 + 0.31, 0.5 - 0.21 + 2 * TWO_PI, 1.9])
    M = _particles(r, fi, np.zeros(6))
    Er, Ez, Efi, gap = _eval(M, gaz, 1.0, r_axis, p_axis, *maps)
    # 0: held by gaps 0 (local 0.25) and 1 (0.15): the first wins; 1: local
    # 0.39 to gap 0 (outside), 0.29 to gap 1; 2: gap 2
    assert gap.tolist() == [0, 1, 2, 0, 0, 0]
    np.testing.assert_allclose(Er[0], fs[0](8.5, 0.25), rtol=1e-12)
    np.testing.assert_allclose(Er[1], fs[0](8.5, 0.29), rtol=1e-12)
    np.testing.assert_allclose(Er[2], fs[0](8.5, 0.1), rtol=1e-12)
    assert Er[3] == 0.0 and Efi[3] == 0.0              # 0.31 beyond gap 2's window, not in 0/1's
    assert Er[4] == 0.0                                # -0.21 below every window
    assert Er[5] == 0.0 and Ez[5] == 0.0 and Efi[5] == 0.0


def test_ez_is_linear_in_z():
    r_axis, p_axis, fs, maps = _grid()
    z = np.linspace(-0.03, 0.03, 7)
    M = _particles(np.full(7, 8.43), np.full(7, 0.1), z)
    _, Ez, _, _ = _eval(M, np.array([0.0]), 2.0, r_axis, p_axis, *maps)
    np.testing.assert_allclose(Ez, 2.0 * fs[1](8.43, 0.1) * z, rtol=1e-12, atol=1e-15)


def test_planted_transposed_weights_are_caught():
    """Bilinear with the r and phi weights exchanged (a plausible slip) misses
    the closed form the exactness test uses."""
    r_axis, p_axis, fs, maps = _grid()
    r, loc = 8.537, 0.123
    dr, dp = r_axis[1] - r_axis[0], p_axis[1] - p_axis[0]
    i, j = int((r - r_axis[0]) // dr), int((loc - p_axis[0]) // dp)
    tx, ty = (r - r_axis[i]) / dr, (loc - p_axis[j]) / dp
    m = maps[0]
    good = (1 - tx) * (1 - ty) * m[i, j] + (1 - tx) * ty * m[i, j + 1] + tx * (1 - ty) * m[i + 1, j] + tx * ty * m[i + 1, j + 1]
    bad = (1 - ty) * (1 - tx) * m[i, j] + (1 - ty) * tx * m[i, j + 1] + ty * (1 - tx) * m[i + 1, j] + ty * tx * m[i + 1, j + 1]
    assert abs(good - fs[0](r, loc)) < 1e-12 and abs(bad - fs[0](r, loc)) > 1e-3


def test_spiral_emap_files_round_trip(tmp_path):
    from pyffa_amd import lattice as Lt
    from pyffa_amd.FFAG_Field import FFAG_EField_spiral
    path = str(tmp_path / "spiral" / "FrequencyCurve_spiral.txt")
    Lt.write_spiral_emap(path, gap_azimuth_deg=(15.0, 195.0), f_start=3.6e6, f_end=3.7e6, n_t=11, gap_width=0.05,
                         nr=31, nphi=41)
    em = FFAG_EField_spiral(path, True)
    assert em.GapType == "spiral" and em.EnableFlag
    np.testing.assert_allclose(em.gap_azimuths, np.deg2rad([15.0, 195.0]))
    np.testing.assert_allclose(em.rf_shift, np.deg2rad([0.0, -180.0]))
    assert em.Er_map.shape == (31, 41) and em.Ezcoef_map.shape == (31, 41) and em.Ephi_map.shape == (31, 41)
    assert em.nr == 31 and em.nphi == 41 and abs(em.acc_Etot_uniform - 20.0) < 1e-12
    assert em.freq_curve.shape == (11, 4)
    f, v = em.Interpolation1D_freq_curve(em.freq_curve[3, 0])
    assert abs(f - em.freq_curve[3, 1]) < 1e-6 and v == 100e3
    # the normal-field profile integrates to gap_width across the line
    d_phi = em.phi_axis[1] - em.phi_axis[0]
    k = em.nr // 2
    line = np.hypot(em.Er_map[k], em.Ephi_map[k])
    integral = line.sum() * d_phi * em.r_axis[k] * np.cos(np.deg2rad(35.0))
    assert abs(integral / 0.05 - 1.0) < 0.02
