"""Pins for oracle functions that the GPU is compared against.

Each check derives its expectation independently of the oracle's code:
from the physics (a relativistic Lorentz force in Cartesian coordinates, a
closed-form energy gain, gamma from a constructed kinetic energy), from
exactness properties (bilinear interpolation reproduces bilinear tables,
including where clamped cells extrapolate), from closed forms (Coulomb
field of two charges, antisymmetry of mirrored charges) or from a schedule
worked out by hand from the reference's rules.  Each test also shows that a
plausible slip (raw vs sector-reduced azimuth, transposed table, a flipped
magnetic term, a wrong extrapolation factor) gives a different answer, so
the check can fail.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

L = O.lib()
p, d, i64, ci = O._p, O._d, O._i64, O._i
NCOL = 16


def _rows(n):
    return np.zeros((n, NCOL))


# ---------------------------------------------------------------------------
# bharmonics_kernel_1d (src/FFAG_Field.py:1541-1587) -> ora_bharmonics
# ---------------------------------------------------------------------------
def _bharm(M, m, amp, ph, rmin, rmax, Bz0, Br0, Bfi0):
    Bz, Br, Bfi = Bz0.copy(), Br0.copy(), Bfi0.copy()
    L.ora_bharmonics(p(M), i64(M.shape[0]), ci(m.size), p(m), p(amp), p(ph), p(rmin), p(rmax), p(Bz), p(Br), p(Bfi))
    return Bz, Br, Bfi


def test_bharmonics_single_modes_window_edges_and_raw_azimuth():
    m = np.array([3, 5], np.int64)
    amp = np.array([1e-3, 2e-4])
    ph = np.array([0.4, -1.0])
    rmin = np.array([8.2, 8.5])
    rmax = np.array([8.8, 9.0])
    r = np.array([8.2, 8.8, 8.2 - 1e-9, 8.8 + 1e-9, 8.6, 8.95, 9.0, 0.0, -1.0])
    fi = np.array([7.3, -2.1, 15.0, 0.3, 40.0, -33.0, 2.5, 1.0, 1.0])   # raw (unwrapped) azimuths
    M = _rows(r.size)
    M[:, 0], M[:, 4] = r, fi
    base = np.linspace(0.1, 0.9, r.size)
    Bz, Br, Bfi = _bharm(M, m, amp, ph, rmin, rmax, base, base + 1, base + 2)
    # Bz += A cos(m phi + phase) inside [rmin, rmax] (both ends inclusive)
    expect = base.copy()
    expect[0] += 1e-3 * np.cos(3 * 7.3 + 0.4)                      # r == rmin of harmonic 0
    expect[1] += 1e-3 * np.cos(3 * -2.1 + 0.4) + 2e-4 * np.cos(5 * -2.1 - 1.0)   # r == rmax of 0, inside 1
    expect[3] += 2e-4 * np.cos(5 * 0.3 - 1.0)                       # just above harmonic 0's window
    expect[4] += 1e-3 * np.cos(3 * 40.0 + 0.4) + 2e-4 * np.cos(5 * 40.0 - 1.0)
    expect[5] += 2e-4 * np.cos(5 * -33.0 - 1.0)
    expect[6] += 2e-4 * np.cos(5 * 2.5 - 1.0)                       # r == rmax of harmonic 1
    # row 2 just below both windows, rows 7, 8 with r <= 0: unchanged
    np.testing.assert_allclose(Bz, expect, rtol=0, atol=1e-15)
    np.testing.assert_array_equal(Br, base + 1)      # the model has no Br / Bfi terms
    np.testing.assert_array_equal(Bfi, base + 2)
    # the azimuth is used raw: reducing it to the 30-degree sector first (a
    # plausible slip) changes rows whose m * (phi - phi mod P) is no multiple of 2 pi
    P = 2 * np.pi / 12
    slip = base.copy()
    slip[0] += 1e-3 * np.cos(3 * np.mod(7.3, P) + 0.4)
    assert abs(slip[0] - Bz[0]) > 1e-5


# ---------------------------------------------------------------------------
# Bilinear_interp_2D_vect_uniform (src/FFAG_MathTools.py:827-936) -> ora_bilinear_uniform
# ---------------------------------------------------------------------------
def _bil(X0, dX, NX, Y0, dY, NY, Z, x, y):
    out = np.zeros(x.size)
    L.ora_bilinear_uniform(d(X0), d(dX), i64(NX), d(Y0), d(dY), i64(NY), p(np.ascontiguousarray(Z)),
                           p(np.ascontiguousarray(x)), p(np.ascontiguousarray(y)), i64(x.size), p(out))
    return out


def test_bilinear_uniform_exact_on_bilinear_tables_and_clamped_extrapolation():
    X0, dX, NX = 280.0, 20.0, 17          # SEO-like Ek axis (MeV)
    Y0, dY, NY = 0.0, 2 * np.pi / 999, 1000
    X = X0 + dX * np.arange(NX)
    Y = Y0 + dY * np.arange(NY)
    f = lambda x, y: 8.5 + 0.003 * x - 0.02 * y + 1e-4 * x * y
    Z = f(X[:, None], Y[None, :])
    rng = np.random.default_rng(1)
    # inside, on nodes, and outside both axes (indices clamp to the edge
    # cells, whose bilinear form extrapolates a bilinear function exactly)
    x = np.concatenate([rng.uniform(X[0], X[-1], 200), X[[0, 5, -1]], [250.0, 700.0, 260.0, 640.0]])
    y = np.concatenate([rng.uniform(Y[0], Y[-1], 200), Y[[0, 7, -1]], [-0.3, 7.0, 6.5, -1.0]])
    got = _bil(X0, dX, NX, Y0, dY, NY, Z, x, y)
    np.testing.assert_allclose(got, f(x, y), rtol=1e-13, atol=1e-12)
    # a table of x^2 beyond the last node: the last cell's chord, extended
    Zq = np.repeat((X ** 2)[:, None], NY, axis=1)
    xq = np.array([X[-1] + 7.0, X[0] - 11.0])
    yq = np.array([1.0, 2.0])
    chord_hi = X[-2] ** 2 + (X[-1] ** 2 - X[-2] ** 2) * (xq[0] - X[-2]) / dX
    chord_lo = X[0] ** 2 + (X[1] ** 2 - X[0] ** 2) * (xq[1] - X[0]) / dX
    np.testing.assert_allclose(_bil(X0, dX, NX, Y0, dY, NY, Zq, xq, yq), [chord_hi, chord_lo], rtol=1e-13)
    # row i is the x index: a transposed read (slip) of a non-symmetric table differs
    sq = np.add.outer(np.arange(5.0), 10 * np.arange(5.0))
    a = _bil(0.0, 1.0, 5, 0.0, 1.0, 5, sq, np.array([1.25]), np.array([3.5]))
    b = _bil(0.0, 1.0, 5, 0.0, 1.0, 5, sq.T.copy(), np.array([1.25]), np.array([3.5]))
    assert a[0] == pytest.approx(1.25 + 35.0) and abs(a[0] - b[0]) > 1.0


# ---------------------------------------------------------------------------
# RK4_equations_of_motion_njit (src/FFAG_MathTools.py:2205-2266) -> ora_rk4_rhs
# ---------------------------------------------------------------------------
def _lorentz_cyl(r, fi, vr, vz, Et, E, B_cyl_ref):
    """Cylindrical rates of (r, vr, z, vz, phi, E_total) from the relativistic
    Lorentz force in Cartesian coordinates, dp/dt = q (E + v x B) with
    a = (q c^2 / E_tot) (E + v x B - v (v . E) / c^2), for the reference's
    field convention B = -(Br, Bphi, Bz) (its Cartesian field is the negated
    rotated map field, src/FFAG_MathTools.py:2011-2014)."""
    q, c, E0 = O.Q, O.C, O.E0_J
    gam = Et / E0
    v = c * np.sqrt(1 - 1 / gam ** 2)
    vf = np.sqrt(v ** 2 - vr ** 2 - vz ** 2)
    er_hat = np.stack([np.cos(fi), np.sin(fi), np.zeros_like(fi)], 1)
    ef_hat = np.stack([-np.sin(fi), np.cos(fi), np.zeros_like(fi)], 1)
    ez_hat = np.tile([0.0, 0.0, 1.0], (fi.size, 1))
    cart = lambda a, b, cz: a[:, None] * er_hat + b[:, None] * ef_hat + cz[:, None] * ez_hat
    vv = cart(vr, vf, vz)
    EE = cart(E[0], E[1], E[2])
    BB = -cart(B_cyl_ref[0], B_cyl_ref[1], B_cyl_ref[2])
    vdotE = (vv * EE).sum(1)
    acc = (q * c * c / Et)[:, None] * (EE + np.cross(vv, BB) - vv * (vdotE / (c * c))[:, None])
    out = np.zeros((fi.size, 6))
    out[:, 0] = vr
    out[:, 1] = (acc * er_hat).sum(1) + vf ** 2 / r       # d(dr/dt)/dt = a_r + v_phi^2 / r
    out[:, 2] = vz
    out[:, 3] = acc[:, 2]
    out[:, 4] = vf / r
    out[:, 5] = q * vdotE
    return out


def _rk4_rhs(M, Er, Ez, Efi, Br, Bz, Bfi, rate=0.0):
    n = M.shape[0]
    dx = np.zeros((n, NCOL))
    gap = np.zeros(n, np.int32)
    L.ora_rk4_rhs(p(M), i64(n), p(Er), p(Ez), p(Efi), p(Br), p(Bz), p(Bfi), d(O.Q), d(O.C), d(O.E0_J), p(dx),
                  d(rate), p(np.zeros(1)), p(gap), ci(0))
    return dx


def test_rk4_rhs_is_the_relativistic_lorentz_force():
    rng = np.random.default_rng(5)
    n = 64
    ek = rng.uniform(250e6, 600e6, n) * O.Q
    Et = ek + O.E0_J
    v = O.C * np.sqrt(1 - (O.E0_J / Et) ** 2)
    r = rng.uniform(8.0, 9.3, n)
    fi = rng.uniform(-10, 10, n)
    vr = 0.05 * v * rng.standard_normal(n)
    vz = 0.02 * v * rng.standard_normal(n)
    E = rng.normal(0, 2e6, (3, n))            # (Er, Ephi, Ez) [V/m], carried in the SC columns
    B = rng.normal(0, 1.5, (3, n))            # map field (Br, Bphi, Bz) [T]
    M = _rows(n)
    M[:, 0], M[:, 1], M[:, 3], M[:, 4], M[:, 5] = r, vr, vz, fi, Et
    M[:, 7] = M[:, 8] = 1.0
    M[:, 10], M[:, 11], M[:, 12] = E[0], E[2], E[1]
    zero = np.zeros(n)
    dx = _rk4_rhs(M, zero, zero, zero, B[0], B[2], B[1], rate=123.0)
    want = _lorentz_cyl(r, fi, vr, vz, Et, E, B)
    got = dx[:, [0, 1, 2, 3, 4, 5]]
    scale = np.abs(want).max(0)
    np.testing.assert_allclose(got / scale, want / scale, rtol=0, atol=1e-13)
    np.testing.assert_array_equal(dx[:, 9], 123.0)
    # a flipped magnetic term (the physical sign of B, a plausible slip) is caught
    slip = _lorentz_cyl(r, fi, vr, vz, Et, E, -B)
    assert np.abs(slip[:, 1] - got[:, 1]).max() / scale[1] > 1e-3
    # not injected / not alive: only the RF-phase rate is written
    M2 = M.copy()
    M2[::2, 7] = 0.0
    M2[1::2, 8] = 0.0
    dx2 = _rk4_rhs(M2, zero, zero, zero, B[0], B[2], B[1], rate=7.0)
    assert np.all(dx2[:, :6] == 0.0) and np.all(dx2[:, 9] == 7.0)


def test_rk4_energy_gain_in_a_uniform_longitudinal_field():
    """rk4_step with no magnetic field and a uniform Ez from the SC columns:
    p_z grows by q Ez t, so E_tot(t) = sqrt(E_perp^2 + c^2 (p_z0 + q Ez t)^2)."""
    mc = O.Machine(r_min=8.0, r_step=0.001, r_size=2, fi_min=0.0, fi_step=0.1, fi_size=2,
                   Bz_stack=np.zeros((1, 2, 2)), Br_stack=np.zeros((1, 2, 2)), Bfi_stack=np.zeros((1, 2, 2)),
                   max_order=0, s_poly=np.zeros(1), sp_poly=np.zeros(1), B0=0.0, k_value=0.0, nsectors=12,
                   convex=np.zeros((0, 5)))
    n = 4
    Et0 = (np.array([300.0, 300.0, 450.0, 600.0]) * 1e6 * O.Q) + O.E0_J
    v = O.C * np.sqrt(1 - (O.E0_J / Et0) ** 2)
    vz0 = np.array([0.0, 0.01, -0.02, 0.005]) * v
    vr0 = np.array([0.0, 0.02, 0.01, -0.03]) * v
    Ez = np.array([5e6, -3e6, 8e6, 1e6])
    M = _rows(n)
    M[:, 0], M[:, 1], M[:, 3], M[:, 4], M[:, 5] = 8.5, vr0, vz0, 0.3, Et0
    M[:, 7] = M[:, 8] = 1.0
    M[:, 11] = Ez
    sc = O.Scratch(n)
    h, steps = 1e-10, 200
    for k in range(steps):
        O.rk4_step(mc, M, sc, k * h, h)
    t = steps * h
    pz0 = (Et0 / O.C ** 2) * vz0
    Eperp2 = Et0 ** 2 - (O.C * pz0) ** 2
    want = np.sqrt(Eperp2 + (O.C * (pz0 + O.Q * Ez * t)) ** 2)
    np.testing.assert_allclose(M[:, 5], want, rtol=1e-12)
    assert np.all(np.abs(M[:, 5] - Et0) > 1e-3 * O.Q * np.abs(Ez) * t)   # the field did work


# ---------------------------------------------------------------------------
# GetLocalCoordinates_2p5(_rk) (src/FFAG_SC_new.py:74-148) -> ora_sc_local_coords
# ---------------------------------------------------------------------------
def _local(M, mask, rk):
    n = M.shape[0]
    out = np.zeros(6 * n)
    sg = ctypes.c_double(0.0)
    cnt = L.ora_sc_local_coords(p(M), i64(n), p(mask.astype(np.uint8)), d(O.C), ci(1 if rk else 0), p(out),
                                ctypes.byref(sg))
    return cnt, [out[j * n: j * n + cnt] for j in range(6)], sg.value


@pytest.mark.parametrize("rk", [False, True])
def test_sc_local_coords_gamma_compaction_and_azimuth(rk):
    rng = np.random.default_rng(9)
    n = 500
    ek_mev = rng.uniform(250.0, 350.0, n)
    gam = 1.0 + ek_mev / 938.2723                      # the kinetic energy each row is built from
    v = O.C * np.sqrt(1 - 1 / gam ** 2)
    a, b = rng.uniform(-0.05, 0.05, n), rng.uniform(-0.03, 0.03, n)
    vr, vz = a * v, b * v
    vf = np.sqrt(v ** 2 - vr ** 2 - vz ** 2)
    r = rng.uniform(8.0, 9.3, n)
    fi = rng.uniform(-30.0, 30.0, n)
    M = _rows(n)
    M[:, 0], M[:, 1], M[:, 2], M[:, 3], M[:, 4] = r, vr, rng.normal(0, 0.01, n), vz, fi
    M[:, 5] = gam * O.E0_J if rk else vf / r
    mask = rng.random(n) < 0.7
    cnt, (R, FI, Z, G, FM, IX), sg = _local(M, mask, rk)
    keep = np.flatnonzero(mask)
    assert cnt == keep.size
    np.testing.assert_array_equal(IX, keep)                          # survivors in row order
    np.testing.assert_array_equal(R, r[keep])
    np.testing.assert_array_equal(FI, fi[keep])
    np.testing.assert_allclose(G, gam[keep], rtol=2e-14 if rk else 1e-12)
    assert np.all((FM >= 0) & (FM < 2 * np.pi))
    np.testing.assert_allclose(np.cos(FM), np.cos(fi[keep]), atol=1e-12)
    np.testing.assert_allclose(np.sin(FM), np.sin(fi[keep]), atol=1e-12)
    assert sg == pytest.approx(gam[keep].sum(), rel=1e-13)


# ---------------------------------------------------------------------------
# total_field_1d_numba (src/FFAG_SC_new.py:763-782) -> ora_total_field_1d
# ---------------------------------------------------------------------------
def _tf1d(s, q):
    E = np.full(s.size, np.nan)
    L.ora_total_field_1d(p(np.ascontiguousarray(s, dtype=np.float64)), p(np.ascontiguousarray(q, dtype=np.float64)),
                         ci(s.size), p(E))
    return E


def test_total_field_1d_coulomb_mirror_and_broadcast_sum():
    K = 1.0 / (4 * np.pi * 8.854187817e-12)
    # two charges: E at each is K q_other sign(s_i - s_j) / (s_i - s_j)^2
    E2 = _tf1d(np.array([0.0, 0.5]), np.array([2e-9, 3e-9]))
    np.testing.assert_allclose(E2, [K * 3e-9 * (-1) / 0.25, K * 2e-9 / 0.25], rtol=1e-14)
    # mirrored charges on a symmetric grid: the field is odd about the centre
    s = np.linspace(-2.0, 2.0, 41)
    rng = np.random.default_rng(3)
    q = rng.uniform(0.5, 2.0, 21) * 1e-10
    q = np.concatenate([q, q[-2::-1]])
    E = _tf1d(s, q)
    np.testing.assert_allclose(E, -E[::-1], rtol=0, atol=1e-12 * np.abs(E).max())
    assert E[20] == pytest.approx(0.0, abs=1e-12 * np.abs(E).max())
    # independent vectorised sum over all pairs
    s = np.sort(rng.uniform(0, 30, 64))
    s = np.unique(s)
    q = rng.uniform(0, 1e-9, s.size)
    dd = s[:, None] - s[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        w = np.where(dd != 0, np.sign(dd) / dd ** 2, 0.0)
    np.testing.assert_allclose(_tf1d(s, q), K * (w @ q), rtol=1e-12)
    # a non-increasing grid returns zeros (the reference's early exit)
    np.testing.assert_array_equal(_tf1d(np.array([1.0, 1.0, 2.0]), np.ones(3)), 0.0)


# ---------------------------------------------------------------------------
# The tracker's SC endpoint schedule (src/FFAG_track.py:331-478) -> oracle.Tracker
# ---------------------------------------------------------------------------
def test_tracker_endpoint_roll_and_extrapolation_schedule(monkeypatch):
    """SC every 4 steps over 12 steps (solves at steps 0, 4, 8); particles 0
    and 1 injected at step 0, particle 2 at step 2, particle 3 at step 6;
    particle 1 lost in step 5.  Each solve k writes E_k(i) to the survivors'
    columns 10-12; the push is replaced by a recorder.  Expected columns
    10-12 seen by each push, from the reference's rules:
      steps 0-3  (s = 0):            p0, p1: E0;      p2: untouched
      steps 4-7  (s = k/4, k=0..3):  p0, p1: E1 + s (E1 - E0) (p1 also after
                                     its loss: masks change only at solves);
                                     p2: E1 (no earlier endpoint); p3 untouched
      steps 8-11 (s = k/4):          p0, p2: E2 + s (E2 - E1); p3: E2;
                                     p1: its step-7 value (no longer written)."""
    n, sc_step, nsteps, dt = 4, 4, 12, 1e-9
    M = _rows(n)
    M[:, 6] = [0.0, 0.0, 2.0, 6.0]            # inj_t [ns]
    M[:, 10:13] = -7.0                        # sentinel of never-written columns
    solves = []
    seen = []

    def E(k, i):
        return np.array([100.0 * (k + 1) + i, 200.0 * (k + 1) + 3 * i, 300.0 * (k + 1) - i])

    def fake_sc_solve(Mx, step_survive, *args, **kw):
        k = len(solves)
        for i in np.flatnonzero(step_survive):
            Mx[i, 10:13] = E(k, i)
        solves.append(np.flatnonzero(step_survive).tolist())
        return {}

    def fake_push(mc, Mx, sc, t, h):
        seen.append(Mx[:, 10:13].copy())
        if len(seen) == 6:                    # during step 5
            Mx[1, 8] = 0.0

    monkeypatch.setattr(O, "sc_solve", fake_sc_solve)
    monkeypatch.setattr(O, "boris_step", fake_push)
    tr = O.Tracker(None, M, enable_sc=True, sc_step=sc_step, grid=(4, 4, 4, 4.0))
    tr.run(0.0, dt, nsteps)
    assert solves == [[0, 1], [0, 1, 2], [0, 2, 3]]
    S = np.array(seen)
    sentinel = np.full(3, -7.0)
    for st in range(nsteps):
        s = (st % sc_step) / sc_step if st >= sc_step else 0.0
        if st < 4:
            want = {0: E(0, 0), 1: E(0, 1), 2: sentinel, 3: sentinel}
        elif st < 8:
            want = {0: E(1, 0) + s * (E(1, 0) - E(0, 0)), 1: E(1, 1) + s * (E(1, 1) - E(0, 1)), 2: E(1, 2),
                    3: sentinel}
        else:
            want = {0: E(2, 0) + s * (E(2, 0) - E(1, 0)), 1: E(1, 1) + 0.75 * (E(1, 1) - E(0, 1)),
                    2: E(2, 2) + s * (E(2, 2) - E(1, 2)), 3: E(2, 3)}
        for i in range(n):
            np.testing.assert_allclose(S[st, i], want[i], rtol=0, atol=1e-12, err_msg=f"step {st} particle {i}")
    # an extrapolation factor without the 'steps_i >= SC_step' guard, or
    # without the roll, would differ at step 9 for particle 0
    assert abs(S[9, 0, 0] - (E(2, 0)[0] + 0.25 * (E(2, 0)[0] - E(0, 0)[0]))) > 1.0
