"""GPU closed-orbit, tune and Twiss finder (ffa_orbit_track, FFAG_SEO)
against the oracle's restatement of the reference's SEO search, the
uniform-field closed forms, and the files the bunch generator reads."""
import math
import os

import numpy as np
import pytest

from oracle import orbit as OB
from tests import cases

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

EK = 300.0


def _gp(harmonics=None):
    mc, gp = cases.machines("stable", max_order=0, harmonics=harmonics)
    return mc, gp


def test_orbit_tracking_matches_the_oracle():
    from pyffa_amd.FFAG_SEO import OrbitTracker
    print("orbit tracking vs oracle", flush=True)
    harm = [dict(m=3, amp_Gs=2.0, phase_deg=30.0, rmin=8.0, rmax=9.3)]
    mc, gp = _gp(harm)
    X = np.array([[8.40, 0.12, 0.0, 0.0], [8.45, 0.10, 2e-3, -1e-3], [8.60, -0.05, -5e-3, 2e-3]])
    ek = np.array([300.0, 310.0, 290.0])
    tr = OrbitTracker(gp)
    tp, traj = tr.track(X, ek, 0.0, np.pi / 6, 0.001, record=True)
    tp_o, traj_o = OB.rk4_solve(mc, X, [OB.Q / OB.ek_to_p(e) for e in ek], 0.0, np.pi / 6, 0.001, record=True)
    assert np.array_equal(tp, tp_o) and traj.shape == traj_o.shape
    for c in range(4):
        scale = np.abs(traj_o[:, :, c]).max()
        assert np.max(np.abs(traj[:, :, c] - traj_o[:, :, c])) <= 1e-12 * scale, c
    # over a full turn the third particle leaves the field region (pf becomes
    # NaN) in both: NaN where the oracle has NaN, equal elsewhere
    end = tr.track(X, ek, 0.0, 2 * np.pi, 0.001)
    end_o = OB.rk4_solve(mc, X, [OB.Q / OB.ek_to_p(e) for e in ek], 0.0, 2 * np.pi, 0.001)
    np.testing.assert_allclose(end, end_o, rtol=1e-10, atol=1e-12, equal_nan=True)


def test_closed_orbits_tunes_and_twiss():
    from pyffa_amd import lattice as Lt
    from pyffa_amd.FFAG_SEO import FFAG_BetaFuncCalc, FFAG_SearchSEO
    print("closed orbits, tunes, twiss", flush=True)
    mc, gp = _gp()
    seo = FFAG_SearchSEO(gp)
    eks = np.array([285.0, 300.0, 315.0])
    r0, pr0, _ = seo.search(eks)
    lat = cases.lattice("stable")
    for i, e in enumerate(eks):
        co = Lt.closed_orbit(lat, e, n_theta=2000)
        assert abs(r0[i] / co["r"][0] - 1) < 3e-7 and abs(pr0[i] - math.sin(math.atan(co["pr"][0]))) < 1e-7
        # closed for the oracle's integrator too
        end = OB.rk4_solve(mc, [[r0[i], pr0[i], 0, 0]], [OB.Q / OB.ek_to_p(e)], 0.0, np.pi / 6, 0.001)[0]
        assert abs(end[0] - r0[i]) < 1e-7 * r0[i]
    qr, qz, f, mr, per = seo.tunes(r0, pr0, eks)
    for i, e in enumerate(eks):
        oq = OB.tunes(mc, r0[i], pr0[i], e)
        np.testing.assert_allclose([qr[i], qz[i], f[i], mr[i], per[i]], oq, rtol=1e-9)
    bf = FFAG_BetaFuncCalc()
    fi0, br, bz, ar, az = bf.twiss(gp, r0, pr0, eks, beta_n=7)
    for i, e in enumerate(eks):
        ot = OB.twiss(mc, r0[i], pr0[i], e, fi0)
        np.testing.assert_allclose(br[i], ot[:, 0], rtol=1e-6)
        np.testing.assert_allclose(bz[i], ot[:, 1], rtol=1e-6)
        np.testing.assert_allclose(ar[i], ot[:, 2], rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(az[i], ot[:, 3], rtol=1e-5, atol=1e-6)
    # as in the reference, the Twiss cells are tracked in the field map alone
    _, gph = _gp([dict(m=3, amp_Gs=2.0, phase_deg=30.0, rmin=8.0, rmax=9.3)])
    assert np.array_equal(FFAG_BetaFuncCalc().twiss(gph, r0, pr0, eks, beta_n=7)[1], br)
    # fields along the two-cell track: Bz of the closed-orbit particle, Br and
    # Bfi of the z-displaced one, against the oracle's field at the same points
    tp, tBz, tBr, tBf = bf.trajectory_fields()
    _, _, tr = bf._last_track
    assert tBz.shape == (eks.size, tp.size) and tp[0] == 0.0 and abs(tp[-1] - np.pi / 3) < 1e-12
    for i in range(eks.size):
        oz, _, _ = OB.fields(mc, tr[:, 5 * i, 0], tp, tr[:, 5 * i, 2])
        _, orr, of = OB.fields(mc, tr[:, 5 * i + 3, 0], tp, tr[:, 5 * i + 3, 2])
        np.testing.assert_allclose(tBz[i], oz, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(tBr[i], orr, rtol=1e-12, atol=1e-22)
        np.testing.assert_allclose(tBf[i], of, rtol=1e-12, atol=1e-22)
        assert np.max(np.abs(tBr[i])) > 0 and np.max(np.abs(tBf[i])) > 0     # z = 1e-7 m: small but not zero


def test_uniform_field_closed_forms_on_the_gpu():
    from pyffa_amd.FFAG_SEO import FFAG_BetaFuncCalc, FFAG_SearchSEO
    k, r_exact = -0.36, 8.5
    _, gp = _gp()
    bm = gp.Bmap
    bm.convex_params = np.array([[-10.0, 10.0, 1.0, 1.0, 1.0]])
    bm.k_value = k
    bm.B0 = OB.ek_to_p(EK) / OB.Q / r_exact ** (k + 1)
    bm.s_rad_polynomial = np.zeros(1)
    bm.s_rad_polynomial_prime = np.zeros(1)
    bm.r_min, bm.r_max = 8.0, 9.3
    seo = FFAG_SearchSEO(gp)
    r0, pr0, r_an = seo.search([EK])
    assert abs(r_an[0] / r_exact - 1) < 1e-9                      # the mean-field radius is exact here
    assert abs(r0[0] / r_exact - 1) < 1e-8 and abs(pr0[0]) < 1e-8
    qr, qz, f, mr, per = seo.tunes(r0, pr0, [EK])
    assert abs(qr[0] - math.sqrt(1 + k)) < 1e-6 and abs(qz[0] - math.sqrt(-k)) < 1e-6
    _, br, bz, ar, az = FFAG_BetaFuncCalc().twiss(gp, r0, pr0, [EK], beta_n=5)
    np.testing.assert_allclose(br, r_exact / math.sqrt(1 + k), rtol=1e-4)
    np.testing.assert_allclose(bz, r_exact / math.sqrt(-k), rtol=1e-4)


def test_seo_files_feed_the_tracker_and_the_bunch_generator(tmp_path):
    from pyffa_amd import FFAG_Distribution as FD
    from pyffa_amd.FFAG_SEO import FFAG_BetaFuncCalc, FFAG_SearchSEO
    print("seo files", flush=True)
    _, gp = _gp()
    gp.Bmap.BmapFoldName = str(tmp_path)
    gp.SEO_SaveFold = "resultsSEO"
    eks = np.arange(285.0, 316.0, 5.0)
    folder = FFAG_SearchSEO(gp).SearchSEOsControllerVect(eks)
    assert folder == os.path.join(str(tmp_path), "resultsSEO")
    v, path, data = FD.LoadSEOParams(os.path.join(folder, "SEO_ini.txt"), 300.0)
    assert data.shape == (eks.size, 17) and os.path.samefile(path, str(tmp_path))      # the Bmap folder
    ref = cases.seo("stable")
    rt = np.loadtxt(os.path.join(folder, "SEO_r.txt"), skiprows=1)
    pt = np.loadtxt(os.path.join(folder, "SEO_pr.txt"), skiprows=1)
    assert np.allclose(rt[1:, 0], eks) and np.allclose(rt[0, 1:], ref.fi_axis)
    assert np.max(np.abs(rt[1:, 1:] - ref.r_matrix)) < 5e-6                   # independent closed-orbit search
    # the lattice module's table is linear in theta between its 1000 samples
    # per sector: its slope carries that interpolation error (~2e-5)
    assert np.max(np.abs(pt[1:, 1:] - np.sin(np.arctan(ref.pr_matrix)))) < 5e-5
    assert np.max(np.abs(pt[1:, 1] - np.sin(np.arctan(ref.pr_matrix[:, 0])))) < 1e-7
    assert abs(v[1] - rt[4, 1]) < 1e-7 and np.all(np.abs(data[:, 2] % 1 - 0.70) < 0.02)      # row 4: 300 MeV
    FFAG_BetaFuncCalc().CalcBetaFunc(folder, eks, gp)
    bz_, br_, az_, ar_ = FD.LoadSEOBetaFunc(folder, 300.0, 0.0)
    assert 0 < br_ < 20 and 0 < bz_ < 30 and abs(ar_) < 5 and abs(az_) < 5
    for name in ("Bz.txt", "Br.txt", "Bf.txt"):
        with open(os.path.join(folder, name)) as f:
            assert f.readline().startswith("unitEk=MeV, unitFi=deg, unitB=T")
        tab = np.loadtxt(os.path.join(folder, name), skiprows=1)
        assert tab.shape[0] == eks.size + 1 and np.array_equal(tab[1:, 0], eks) and tab[0, 1] == 0.0
    bzt = np.loadtxt(os.path.join(folder, "Bz.txt"), skiprows=1)
    assert np.all(np.abs(bzt[1:, 1:]) < 5.0) and np.ptp(bzt[1:, 1:]) > 0.1     # the F/D pattern along the orbit
