"""CPU pins of the closed-orbit / tune / Twiss oracle: closed forms of a
uniform-flutter field, the lattice module's independent closed-orbit
search, and FFT tunes of the time-stepped 6D oracle tracker; a planted sign
slip in the Bfi terms fails the tune check."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from oracle import orbit as OB
from tests import cases

EK = 300.0


def _uniform_machine(k=-0.36, r0=8.5):
    """One pole reaching far beyond the sector on both sides: F = 1 exactly,
    so Bz = B0 r^k with no flutter."""
    B0 = OB.ek_to_p(EK) / OB.Q / r0 ** (k + 1)
    z = np.zeros((1, 2, 2))
    return O.Machine(r_min=7.0, r_step=1.0, r_size=2, fi_min=0.0, fi_step=1.0, fi_size=2, Bz_stack=z, Br_stack=z,
                     Bfi_stack=z, max_order=0, s_poly=np.zeros(1), sp_poly=np.zeros(1), B0=B0, k_value=k, nsectors=12,
                     convex=np.array([[-10.0, 10.0, 1.0, 1.0, 1.0]])), r0


def test_uniform_field_closed_forms():
    k = -0.36
    mc, r_exact = _uniform_machine(k)
    Bz, _, _ = OB.fields(mc, np.array([8.0, 9.0]), np.array([0.1, 0.4]), np.zeros(2))
    np.testing.assert_allclose(Bz, mc.B0 * np.array([8.0, 9.0]) ** k, rtol=1e-14)
    r0, pr0 = OB.closed_orbit(mc, EK, r_exact * (1 + 2e-6))      # near the mean-field radius, as the reference starts
    # the search stops once one sector moves r by less than 1e-8 r, i.e. within
    # ~1e-8 / (1 - cos(sector phase advance)) of the fixed point
    assert abs(r0 / r_exact - 1) < 2e-7 and abs(pr0) < 1e-7
    qr, qz, f, mean_r, perim = OB.tunes(mc, r0, pr0, EK)
    assert abs(qr - math.sqrt(1 + k)) < 1e-6 and abs(qz - math.sqrt(-k)) < 1e-6
    assert abs(perim / (2 * np.pi * r_exact) - 1) < 1e-4          # sum over a sector's grid points
    assert abs(mean_r / r_exact - 1) < 2e-7
    tw = OB.twiss(mc, r0, pr0, EK, np.array([0.05, 0.3]))
    np.testing.assert_allclose(tw[:, 0], r_exact / math.sqrt(1 + k), rtol=1e-4)
    np.testing.assert_allclose(tw[:, 1], r_exact / math.sqrt(-k), rtol=1e-4)
    assert np.max(np.abs(tw[:, 2:])) < 1e-3


def test_closed_orbit_matches_the_lattice_module():
    from pyffa_amd import lattice as Lt
    mc, _ = cases.machines("stable", max_order=0)
    lat = cases.lattice("stable")
    co = Lt.closed_orbit(lat, EK, n_theta=2000)
    r0, pr0 = OB.closed_orbit(mc, EK, float(co["r"][0]) + 0.01)
    assert abs(r0 / co["r"][0] - 1) < 2e-7
    # the lattice module's pr is tan(angle); the SEO pr is the direction cosine sin(angle)
    assert abs(pr0 - math.sin(math.atan(co["pr"][0]))) < 1e-8


def _fft_tunes(mc, r0, pr0, turns=64, per_turn=6000):
    """Fractional tunes of the oracle's time-stepped 6D RK4 tracker: a
    particle on the closed orbit and two displaced by 0.2 mm in r and z,
    sampled once per revolution (the period from GetFreq, corrected by the
    reference particle's first turn).  Steps of T / 6000 (~0.05 ns, the
    tracking step) resolve the pole fringes; the deviations from the
    reference particle are the betatron signals."""
    from pyffa_amd import diagnostics as D
    from pyffa_amd.FFAG_ParasAndConversion import FFAG_ConversionTools
    _, _, f_rev, _, _ = OB.tunes(mc, r0, pr0, EK)

    def run(T, n_turns):
        P = np.zeros((3, 16))
        P[:, 0] = r0
        P[:, 1] = math.asin(pr0)                       # the tracker's pr is the angle
        P[1, 0] += 2e-4
        P[2, 2] = 2e-4
        P[:, 5] = EK
        P[:, 7] = P[:, 8] = 1.0
        M = np.ascontiguousarray(FFAG_ConversionTools().ConvertPrzek2Vrzek(P))
        rec = []
        for _ in range(n_turns):
            O.run_c(mc, M, 0.0, T / per_turn, per_turn, rk4=True, head=3)
            rec.append(M[:, :5].copy())
        return np.array(rec)
    T = 1.0 / f_rev
    T = T * (1 - (run(T, 1)[0, 0, 4] - 2 * np.pi) / (2 * np.pi))
    rec = run(T, turns)
    assert abs(rec[-1, 0, 0] - r0) < 1e-6                 # the closed orbit is closed for the 6D tracker too
    x, xp = rec[:, 1, 0] - rec[:, 0, 0], rec[:, 1, 1] - rec[:, 0, 1]
    z, zp = rec[:, 2, 2], rec[:, 2, 3]
    return D.tune_from_xpx(x, xp), D.tune_from_xpx(z, zp)


def _close(q, v):
    f = q % 1.0
    return min(abs(f - v), abs(1 - f - v))


_FFT = {}


@pytest.mark.parametrize("planted", [False, True])
def test_tunes_match_time_domain_tracking(planted):
    mc, _ = cases.machines("stable", max_order=0)
    co_r = cases.seo("stable").r_matrix[3, 0]
    r0, pr0 = OB.closed_orbit(mc, EK, co_r + 0.01)
    kw = dict(bfi_sign=-1.0) if planted else {}
    qr, qz, _, _, _ = OB.tunes(mc, r0, pr0, EK, **kw)
    if (r0, pr0) not in _FFT:
        _FFT[(r0, pr0)] = _fft_tunes(mc, r0, pr0)
    vx, vz = _FFT[(r0, pr0)]
    ok = _close(qr, vx) < 2e-4 and _close(qz, vz) < 2e-4
    assert ok != planted, (qr, qz, vx, vz)
