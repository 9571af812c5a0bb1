"""CPU tests of the RF bucket behind the "match" longitudinal distribution
(src/FFAG_bucket.py) and of the injection point on a centre line
(src/FFAG_Distribution.py:1347-1414).

The bucket is pinned by what its equations fix: the fixed points, the small
oscillation period 2 pi / sqrt(-A'(0) B'(0)), conservation of the grid
Hamiltonian along the RK4 trajectories (a flipped sign in either equation
breaks it), and the sampler's H threshold and projected bunch length.  The
centre-line point is pinned by geometry (a straight line through the origin
crosses the orbit at its own azimuth)."""
import numpy as np
import pytest

from pyffa_amd import FFAG_Distribution as FD
from pyffa_amd.FFAG_bucket import FFAG_bucket
from tests.test_distribution import _seo_folder

# the fixture's revolution period: T(Ek) = 2.8e-7 s - 1e-10 s/MeV (Ek - 300 MeV)
E0, U0, PHI_S = 300.0, 1e5, 30.0


@pytest.fixture(scope="module")
def bucket(tmp_path_factory):
    d = _seo_folder(tmp_path_factory.mktemp("bk"))
    return FFAG_bucket(str(d.parent), Ek0=E0, fi0=PHI_S, U0=U0, step_t=20e-9, step_N=18000)


def test_equations_and_fixed_points(bucket):
    Tc = 2.8e-7
    w0 = 2 * np.pi / Tc
    assert bucket.Tc == pytest.approx(Tc, rel=1e-12) and bucket.w0 == pytest.approx(w0, rel=1e-12)
    assert bucket.zd_u == pytest.approx((np.pi - 2 * np.deg2rad(PHI_S)) / w0, rel=1e-12)
    p, q = np.array([0.7, -1.3]), np.array([3e-9, -20e-9])
    d = bucket.dynamic(0.0, np.column_stack((q, p)))
    # higher energy: shorter period (below transition), so the particle arrives earlier
    np.testing.assert_allclose(d[:, 0], -(1e-10 * p) / Tc, rtol=1e-9)
    np.testing.assert_allclose(d[:, 1], U0 * (np.sin(np.deg2rad(PHI_S) + w0 * q) - np.sin(np.deg2rad(PHI_S))) / Tc / 1e6,
                               rtol=1e-12)
    np.testing.assert_allclose(bucket.dynamic(0.0, np.array([[0.0, 0.0], [bucket.zd_u, 0.0]])), 0.0, atol=1e-6)


def test_small_oscillation_period(bucket):
    Tc, w0 = bucket.Tc, bucket.w0
    a1 = -(-1e-10) / Tc * -1.0                     # A'(0) = T'(E0) / Tc
    b1 = U0 * w0 * np.cos(np.deg2rad(PHI_S)) / Tc / 1e6
    period = 2 * np.pi / np.sqrt(-a1 * b1)
    tr = bucket.rk4_solve_dt(bucket.dynamic, 0.0, np.array([[1e-9, 0.0]]))
    zd = tr[:, 0, 1]
    up = np.nonzero((zd[:-1] < 0) & (zd[1:] >= 0))[0]
    assert up.size >= 2
    measured = np.mean(np.diff(up)) * bucket.step_t
    assert measured == pytest.approx(period, rel=5e-3)


def test_hamiltonian_is_conserved_along_bounded_trajectories(bucket):
    bucket.build_H()
    assert bucket.H_sep < 0.0 < bucket.H_map.max() + 1e-30 or bucket.H_sep < 0.0
    n_sep = bucket.X_sep.shape[0]
    assert bucket.stable_fwd[n_sep:].sum() >= 10          # most of the centre line is inside the bucket
    span = abs(bucket.H_sep)
    for pid in np.nonzero(bucket.stable_fwd)[0]:
        h = bucket.H_at(bucket.tr_fwd[:, pid, 1], bucket.tr_fwd[:, pid, 2])
        assert np.ptp(h) < 2e-3 * span, pid
    # the trajectories that leave the bucket start outside it
    for pid in np.nonzero(~bucket.stable_fwd[n_sep:])[0] + n_sep:
        assert bucket.H_at(bucket.X_in[pid, 0], bucket.X_in[pid, 1]) < bucket.H_sep + 1e-3 * span


def test_scatter_filter_keeps_the_bucket_core_of_the_target_length(bucket):
    L = 80.0
    t_ns, ek, h_thr = bucket.scatter_filter_by_span(20_000, L, show=False, rng=np.random.default_rng(3))
    assert t_ns.shape == (20_000,) and ek.shape == (20_000,)
    assert bucket.H_sep < h_thr < 0.0
    h = bucket.H_at(t_ns * 1e-9, ek - E0)
    assert h.min() >= h_thr
    assert np.ptp(t_ns) == pytest.approx(L, rel=0.03)
    # a larger target bunch length means a lower threshold (a larger region)
    _, _, h2 = bucket.scatter_filter_by_span(100, 120.0, show=False, rng=np.random.default_rng(4))
    assert h2 < h_thr


def test_matched_sub_bunch(tmp_path):
    d = _seo_folder(tmp_path)
    FD.set_seed(9)
    M, _, centre, bk = FD.GenerateInitDistributionDtInjNew(str(d), 15.0, 300.0, [2.0, 1.0, 1.0], 5000, 7.0, 0.0,
                                                           40.0, 60.0, acc_U0=U0, acc_phi0=PHI_S,
                                                           dist_types=("gauss", "match"))
    assert isinstance(bk, FFAG_bucket) and M.shape == (5000, 16)
    t = M[:, 6] - 7.0
    assert np.ptp(t) == pytest.approx(60.0, rel=0.05)
    h = bk.H_at(t * 1e-9, M[:, 5] - 300.0)
    assert h.min() >= bk.H_sep


def test_centre_line_crossing_point():
    ek = np.array([280.0, 290.0, 300.0, 310.0, 320.0])
    fi = np.linspace(0, 2 * np.pi, 1201)
    r = 8.0 + 0.01 * ek[:, None] + 0.02 * np.cos(12 * fi)[None, :]
    pr = 0.05 * np.sin(12 * fi)[None, :] + 0 * ek[:, None]
    tab = lambda v: np.block([[np.zeros((1, 1)), fi[None, :]], [ek[:, None], v]])
    a = np.deg2rad(16.1)
    cl = np.array([[9.0 * np.cos(a), 9.0 * np.sin(a)], [10.0 * np.cos(a), 10.0 * np.sin(a)],
                   [12.0 * np.cos(a), 12.0 * np.sin(a)]])
    th, rc, prc = FD.GetInitRAndFiForInjection(tab(r), tab(pr), cl, 305.0)
    assert th == pytest.approx(a, abs=1e-12)
    assert rc == pytest.approx(8.0 + 3.05 + 0.02 * np.cos(12 * a), abs=6e-5)     # chord of a 0.3 deg segment
    k = int(np.floor(a / (2 * np.pi / 1200)))                                    # the segment's first node
    assert prc == pytest.approx(pr[0, k], abs=1e-15)
    with pytest.raises(ValueError):
        FD.GetInitRAndFiForInjection(tab(r), tab(pr), cl * 0.5, 305.0)            # never reaches the orbit
