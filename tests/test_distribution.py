"""CPU tests of the bunch generator: each distribution's rms emittance and
Twiss parameters from its moments, the KV / water-bag boundaries in
normalised phase space, the SEO and Twiss table lookups on closed forms, and
GenerateBunches' painting offsets, ids and injection times."""
import os

import numpy as np
import pytest

from pyffa_amd import FFAG_Distribution as FD

EPS, ALPHA, BETA = 2.0e-6, -1.2, 3.0


def _twiss_from_moments(x, xp):
    dx, dp = x - x.mean(), xp - xp.mean()
    e = np.sqrt(np.mean(dx * dx) * np.mean(dp * dp) - np.mean(dx * dp) ** 2)
    return e, -np.mean(dx * dp) / e, np.mean(dx * dx) / e


@pytest.mark.parametrize("kind", ["gauss", "kv", "waterbag"])
def test_2d_distributions_have_the_rms_emittance_and_twiss(kind):
    FD.set_seed(1)
    tw = (FD.TwissContainer(ALPHA, BETA, EPS), FD.TwissContainer(0.4, 1.5, 3 * EPS))
    x, xp, y, yp = FD.create_distribution(kind, twiss_transverse=tw).sample(400_000)
    for (u, up), (a, b, e) in (((x, xp), (ALPHA, BETA, EPS)), ((y, yp), (0.4, 1.5, 3 * EPS))):
        e_m, a_m, b_m = _twiss_from_moments(u, up)
        assert abs(e_m / e - 1) < 0.01 and abs(a_m - a) < 0.02 and abs(b_m / b - 1) < 0.01, (kind, e_m, a_m, b_m)


@pytest.mark.parametrize("kind", ["gauss", "kv", "waterbag"])
def test_1d_distributions_have_the_rms_emittance(kind):
    FD.set_seed(2)
    u, up = FD.create_distribution(kind, twiss_longitudinal=FD.TwissContainer(ALPHA, BETA, EPS)).sample(400_000)
    e_m, a_m, b_m = _twiss_from_moments(u, up)
    assert abs(e_m / EPS - 1) < 0.01 and abs(a_m - ALPHA) < 0.02 and abs(b_m / BETA - 1) < 0.01


def test_kv_and_waterbag_boundaries():
    FD.set_seed(3)
    tw = (FD.TwissContainer(ALPHA, BETA, EPS), FD.TwissContainer(0.3, 2.0, EPS))
    kv = FD.KVDist2D(*tw)
    x, xp, y, yp = kv.sample(50_000)
    h = kv.twissX.getNormalizedH(x, xp) + kv.twissY.getNormalizedH(y, yp)
    np.testing.assert_allclose(h, 1.0, atol=1e-12)               # on the 4 eps hyper-ellipsoid surface
    wb = FD.WaterBagDist2D(*tw)
    x, xp, y, yp = wb.sample(200_000)
    cx, cy = wb.getTwissContainers()
    h = cx.getNormalizedH(x, xp) + cy.getNormalizedH(y, yp)
    assert h.max() <= 1.0 + 1e-12
    assert abs(np.median(h) - np.sqrt(0.5)) < 0.005             # uniform in a 4-ball: P(h <= s) = s^2
    kv1 = FD.KVDist1D(FD.TwissContainer(ALPHA, BETA, EPS))
    u, up = kv1.sample(10_000)
    np.testing.assert_allclose(kv1.twiss.getNormalizedH(u, up), 1.0, atol=1e-12)
    assert kv1.twiss.emittance == 2 * EPS


def test_gauss_cutoff_and_hollow_waterbag_limits():
    FD.set_seed(4)
    g = FD.GaussDist1D(FD.TwissContainer(ALPHA, BETA, EPS), cut_off=2.0)
    u, up = g.sample(100_000)
    assert g.twiss.getNormalizedH(u, up).max() <= 4.0 + 1e-12
    hw = FD.IndependentHollowWaterBagDist2D(FD.TwissContainer(0.0, 2.0, EPS), FD.TwissContainer(0.0, 1.0, EPS))
    x, xp, y, yp = hw.sample(20_000)
    q = x ** 2 / hw.hollow_x.kv_dist.twiss.getU_Max() ** 2 + y ** 2 / hw.hollow_y.kv_dist.twiss.getU_Max() ** 2
    assert x.size == 20_000 and q.min() >= 0.3 and q.max() <= 1.0


def test_lagrange_2d_is_exact_on_cubics_including_the_edges():
    X = np.array([280.0, 285.0, 291.0, 300.0, 304.0, 310.0])
    Y = np.linspace(0, 2 * np.pi, 40) ** 1.1
    f = lambda x, y: (1 + 0.2 * x - 1e-3 * x ** 2 + 2e-6 * x ** 3) * (0.5 - y + 0.3 * y ** 2 - 0.02 * y ** 3)
    Z = f(X[:, None], Y[None, :])
    rng = np.random.default_rng(0)
    xi, yi = rng.uniform(275, 315, 200), rng.uniform(-0.2, Y[-1] + 0.2, 200)
    np.testing.assert_allclose(FD.lagrange_2d(X, Y, Z, xi, yi), f(xi, yi), rtol=1e-9, atol=1e-9)


def _seo_folder(tmp_path, ek=(280.0, 290.0, 300.0, 310.0, 320.0)):
    ek = np.asarray(ek)
    info = np.zeros((ek.size, 9))
    info[:, 1] = ek
    info[:, 4] = 1.0 / (2.8e-7 - 1e-10 * (ek - 300))
    info[:, 6] = 8.0 + 0.01 * ek
    info[:, 7] = 1e-3 * (ek - 300)
    info[:, 8] = 53.0 + 0.01 * ek
    d = tmp_path / "Bmap" / "resultsSEO"
    os.makedirs(d)
    np.savetxt(d / "SEO_ini.txt", info, header="a\nb")
    fi = np.linspace(0, 2 * np.pi, 1201)
    for name, tab in (("SEO_r.txt", 8.0 + 0.01 * ek[:, None] + 0.02 * np.cos(12 * fi)[None, :]),
                      ("SEO_pr.txt", 0.05 * np.sin(12 * fi)[None, :] + 0 * ek[:, None])):
        m = np.zeros((ek.size + 1, fi.size + 1))
        m[0, 1:], m[1:, 0], m[1:, 1:] = fi, ek, tab
        np.savetxt(d / name, m, header="h")
    from pyffa_amd import lattice as Lt
    Lt.write_twiss(str(d), Lt.SEOTables(ek, fi, None, None, None, None), 2.5, 3.5, -0.5, 0.25)
    return d


def test_load_seo_params_interpolates_and_extrapolates_linearly(tmp_path):
    d = _seo_folder(tmp_path)
    for e in (300.0, 290.0, 330.0, 270.0):
        v, path, data = FD.LoadSEOParams(str(d / "SEO_ini.txt"), e)
        assert abs(v[1] - (8.0 + 0.01 * e)) < 1e-12 and abs(v[2] - 1e-3 * (e - 300)) < 1e-15
        assert abs(v[3] - (2.8e-7 - 1e-10 * (e - 300))) < 1e-18
    assert os.path.samefile(path, tmp_path / "Bmap") and data.shape == (5, 9)     # the Bmap folder
    assert FD.LoadSEOBetaFunc(str(d), 300.0, 0.3) == pytest.approx((3.5, 2.5, 0.25, -0.5))


def test_generate_bunches_painting(tmp_path):
    d = _seo_folder(tmp_path)
    curve = tmp_path / "curve.paint"
    np.savetxt(curve, np.array([[0.0, 0.0, 0.0, 0.0, 0.0], [10.0, 0.01, 1e-3, 0.002, -1e-3]]), header="t x px y py")
    emap = tmp_path / "emap.txt"
    emap.write_text("gap_azimuth = 0.0\nacc_phi_paint = 30.0\nacc_voltage_paint = 1e5\n#Time (s) f V Ek\n0 1 1 1\n")
    bp = dict(SEO=str(d), BmapPATHName=str(tmp_path / "Bmap"), InjectEk=300.0, InjPosition=15.0, ParticleNum=20_000,
              TransverseREmit=2.0, TransverseZEmit=1.0, LongitudeT=40.0, LongitudeDEk=1.0, InjTimeNanoSec=0,
              TransverseDistType="kv", LongitudeDistType="UniformGauss", PlotFlag=False, PlotRMS=(4, 4, 4),
              PaintEnable=True, PaintMaxNum=4, PaintTimeInterval=3.0, PaintCurve=str(curve), EmapPATHName=str(emap))
    FD.set_seed(5)
    M = FD.GenerateBunches(bp)
    n_sub = 3                                                    # min(floor(10 / 3), 4)
    assert M.shape == (n_sub * 20_000, 16)
    r_c = 8.0 + 0.01 * 300 + 0.02 * np.cos(12 * np.deg2rad(15.0))
    for i in range(n_sub):
        S = M[M[:, 13] == i]
        t = 3.0 * i
        assert S.shape[0] == 20_000 and np.array_equal(S[:, 14], np.arange(20_000))
        assert abs(S[:, 6].mean() - t) < 0.3 and S[:, 6].min() >= t - 20 and S[:, 6].max() <= t + 20
        assert abs(S[:, 0].mean() - (r_c + 0.001 * t)) < 5e-5
        assert abs(S[:, 2].mean() - 0.0002 * t) < 5e-5
        assert np.allclose(S[:, 4], np.deg2rad(15.0))
        e_r, _, _ = _twiss_from_moments(S[:, 0], S[:, 1])
        assert abs(e_r / 2e-6 - 1) < 0.03
        assert abs(S[:, 5].std() - np.sqrt(40.0 / 16 / 40.0)) < 0.01          # sqrt(eps_l / beta_l) MeV
    # an injection centre line file: a straight line through the orbit at 16.1 deg
    cl = tmp_path / "centre_line.cl"
    a = np.deg2rad(16.1)
    np.savetxt(cl, np.array([[9.0 * np.cos(a), 9.0 * np.sin(a)], [12.0 * np.cos(a), 12.0 * np.sin(a)]]), header="x y")
    bp["PaintEnable"], bp["InjPosition"] = False, str(cl)
    M = FD.GenerateBunches(bp)
    assert M.shape == (20_000, 16)
    np.testing.assert_allclose(M[:, 4], a, rtol=1e-12)
