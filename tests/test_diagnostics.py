"""CPU tests of the beam diagnostics: the FFT tune estimator on signals of
known tune, rms emittance on a Twiss ellipse (exact value eps / 2), the SEO
bilinear lookup on affine tables, and the PositionDump file round trip."""
import types

import numpy as np

from pyffa_amd import diagnostics as D


def _tbt(q, n, amp=2e-3, scale=3.0, phase=0.3, offset=0.0):
    k = np.arange(n)
    th = 2 * np.pi * q * k + phase
    return offset + amp * np.cos(th), -amp * scale * np.sin(th)


def test_tune_recovered_for_known_tunes():
    for q in (0.0613, 0.2345, 0.3871, 0.4902):
        x, px = _tbt(q, 512)
        v = D.tune_from_xpx(x, px)
        qx, _ = D.fractional_tunes(v, 0.9)
        assert abs(qx - q) < 2e-5, (q, qx)


def test_tune_ignores_offset_and_amplitude_scale():
    x, px = _tbt(0.2011, 256, amp=1e-4, scale=50.0, offset=5e-2)
    qx, _ = D.fractional_tunes(D.tune_from_xpx(x, px), 0.9)
    assert abs(qx - 0.2011) < 5e-5


def test_parabolic_refinement_beats_the_bin_width():
    """The estimate is far closer than the padded bin width 1 / (8 N): a
    dropped refinement (argmax bin alone) fails this bound."""
    q, n = 0.23456, 256
    x, px = _tbt(q, n)
    qx, _ = D.fractional_tunes(D.tune_from_xpx(x, px), 0.9)
    assert abs(qx - q) < 0.1 / (8 * n)


def test_vertical_fold():
    qx, qz = D.fractional_tunes(np.array([0.7]), np.array([0.6]))
    assert np.allclose(qx, 0.3) and np.allclose(qz, 0.4 - 1.0)
    _, qz = D.fractional_tunes(0.7, 0.8)
    assert np.allclose(qz, 0.2)


def test_rms_emittance_of_twiss_ellipse_is_half_eps():
    eps, alpha, beta = 3.7e-6, -1.3, 2.2
    psi = np.linspace(0, 2 * np.pi, 1000, endpoint=False)
    x = np.sqrt(eps * beta) * np.cos(psi)
    xp = -np.sqrt(eps / beta) * (alpha * np.cos(psi) + np.sin(psi))
    assert abs(D.rms_emittance(x + 0.01, xp - 0.002) / (eps / 2) - 1) < 1e-12


def test_seo_bilinear_exact_on_affine_tables_with_clamped_extrapolation():
    ek = np.linspace(280.0, 320.0, 9)
    fi = np.linspace(0.0, 2 * np.pi, 101)
    E, F = np.meshgrid(ek, fi, indexing="ij")
    f = lambda e, p: 8.0 + 0.01 * e - 0.2 * p + 1e-3 * e * p
    rng = np.random.default_rng(1)
    e_q = rng.uniform(270.0, 330.0, 400)
    p_q = rng.uniform(-0.1, 2 * np.pi + 0.1, 400)
    val, out = D.seo_bilinear(ek, fi, f(E, F), e_q, p_q)
    np.testing.assert_allclose(val, f(e_q, p_q), rtol=1e-13, atol=1e-12)
    assert np.array_equal(out, (e_q < 280) | (e_q >= 320) | (p_q < 0) | (p_q >= 2 * np.pi))


def test_position_dump_round_trip_and_particle_tunes(tmp_path):
    ek = np.linspace(290.0, 310.0, 5)
    fi = np.linspace(0.0, 2 * np.pi, 63)
    seo = types.SimpleNamespace(Ek_axis=ek, fi_axis=fi, r_matrix=np.full((5, 63), 8.5),
                                pr_matrix=np.full((5, 63), 0.01))
    want = {3: (0.21, 0.17), 11: (0.33, 0.12)}
    for gid, (qx, qz) in want.items():
        n = 128
        rows = np.zeros((n, 17))
        x, px = _tbt(qx, n)
        z, pz = _tbt(qz, n, amp=1e-3, scale=1.5)
        rows[:, 0], rows[:, 1], rows[:, 2], rows[:, 3] = 8.5 + x, 0.01 + px, z, pz
        rows[:, 4] = 2 * np.pi * np.arange(n)
        rows[:, 5] = 300.0
        rows[:, 16] = np.arange(n) * 2.8e-7
        perm = np.random.default_rng(gid).permutation(n)      # file order need not be time order
        np.savez(tmp_path / f"particle_{gid}_angle_0.0_rank_0.npz", particles=rows[perm])
    np.savez(tmp_path / "particle_5_angle_90.0_rank_0.npz", particles=np.zeros((3, 17)))
    tbt = D.load_position_dump(str(tmp_path), 0.0)
    assert sorted(tbt) == [3, 11]
    ids, qx, qz = D.particle_tunes(tbt, seo)
    assert ids.tolist() == [3, 11]
    np.testing.assert_allclose(qx, [0.21, 0.33], atol=5e-5)
    np.testing.assert_allclose(qz, [0.17, 0.12], atol=5e-5)
    ids, _, _ = D.particle_tunes(tbt, seo, turns=200)        # too few records: skipped
    assert ids.size == 0
