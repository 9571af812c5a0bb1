"""Entry/exit state conversions (src/FFAG_ParasAndConversion.py:202-333).

The conversions build every tracking input, so they are pinned by
relativity (speed from kinetic energy via p c = sqrt(T^2 + 2 T E0)), by the
reference's definition of pr, pz as angles (tan pr = vr / v_phi,
tan pz = vz / v_phi) and by the round trip through the inverse conversions.
"""
import numpy as np
import pytest

from pyffa_amd.FFAG_ParasAndConversion import FFAG_ConversionTools

Q, C, E0_MEV = 1.60217662e-19, 2.99792458e8, 938.2723


def _points(n=400, seed=0):
    rng = np.random.default_rng(seed)
    P = np.zeros((n, 16))
    P[:, 0] = rng.uniform(8.0, 9.3, n)
    P[:, 1] = rng.uniform(-0.2, 0.2, n)           # pr [rad]
    P[:, 2] = rng.normal(0, 0.01, n)
    P[:, 3] = rng.uniform(-0.1, 0.1, n)           # pz [rad]
    P[:, 4] = rng.uniform(-20, 20, n)
    P[:, 5] = rng.uniform(100.0, 1000.0, n)       # Ek [MeV]
    P[:, 6:] = rng.normal(size=(n, 10))           # inj_t, flags, RF phase, Esc, ids: carried through
    return P


def _speed(ek_mev):
    pc = np.sqrt(ek_mev ** 2 + 2 * ek_mev * E0_MEV)     # MeV
    return C * pc / (ek_mev + E0_MEV)


@pytest.mark.parametrize("boris", [True, False])
def test_przek_to_vrzek_closed_forms(boris):
    ct = FFAG_ConversionTools()
    P = _points()
    V = ct.ConvertPrzek2Vrzek_boris(P) if boris else ct.ConvertPrzek2Vrzek(P)
    v = _speed(P[:, 5])
    vr, vz = V[:, 1], V[:, 3]
    if boris:
        vf = V[:, 5] * V[:, 0]
    else:
        np.testing.assert_allclose(V[:, 5], (P[:, 5] + E0_MEV) * 1e6 * Q, rtol=1e-15)   # E_total [J]
        vf = np.sqrt(v ** 2 - vr ** 2 - vz ** 2)
    np.testing.assert_allclose(np.sqrt(vr ** 2 + vz ** 2 + vf ** 2), v, rtol=1e-13)
    np.testing.assert_allclose(vr / vf, np.tan(P[:, 1]), rtol=1e-13, atol=1e-15)       # tan pr = vr / v_phi
    np.testing.assert_allclose(vz / vf, np.tan(P[:, 3]), rtol=1e-13, atol=1e-15)
    keep = [0, 2, 4] + list(range(6, 16))
    np.testing.assert_array_equal(V[:, keep], P[:, keep])
    # reading pr as a direction cosine instead of an angle (a plausible slip) differs
    assert np.abs(vr / v - np.sin(P[:, 1])).max() > 1e-4


def test_on_orbit_particle():
    """pr = pz = 0: v_r = v_z = 0 and dphi/dt = v / r; 300 MeV protons move at
    beta = 0.65...: sqrt(1 - (E0 / (E0 + T))^2)."""
    ct = FFAG_ConversionTools()
    P = np.zeros((1, 16))
    P[0, 0], P[0, 5] = 8.5, 300.0
    V = ct.ConvertPrzek2Vrzek_boris(P)
    beta = np.sqrt(1 - (E0_MEV / (E0_MEV + 300.0)) ** 2)
    assert V[0, 1] == 0.0 and V[0, 3] == 0.0
    assert V[0, 5] == pytest.approx(beta * C / 8.5, rel=1e-14)


@pytest.mark.parametrize("boris", [True, False])
def test_round_trip(boris):
    ct = FFAG_ConversionTools()
    P = _points(seed=3)
    if boris:
        back = ct.ConvertVrzek2Przek_boris(ct.ConvertPrzek2Vrzek_boris(P))
    else:
        back = ct.ConvertVrzek2Przek(ct.ConvertPrzek2Vrzek(P))
    np.testing.assert_allclose(back[:, [1, 3]], P[:, [1, 3]], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(back[:, 5], P[:, 5], rtol=1e-9)     # Ek through gamma - 1
    keep = [0, 2, 4] + list(range(6, 16))
    np.testing.assert_array_equal(back[:, keep], P[:, keep])


def test_unit_helpers():
    ct = FFAG_ConversionTools()
    ek = np.array([1.0, 300.0, 600.0, 5000.0])
    np.testing.assert_allclose(ct.Ek2v(ek), _speed(ek), rtol=1e-13)
    pc = np.sqrt(ek ** 2 + 2 * ek * E0_MEV) * 1e6 * Q           # J
    np.testing.assert_allclose(ct.Ek2P(ek), pc / C, rtol=1e-13)
    np.testing.assert_allclose(ct.v2Ek(_speed(ek)), ek, rtol=1e-9)
    np.testing.assert_allclose(ct.EtotJ2EkMeV(ct.EkMeV2EtotJ(ek)), ek, rtol=1e-12)
