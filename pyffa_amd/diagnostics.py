"""Beam diagnostics over dumped or device state: closed-orbit coordinates,
rms emittance and the FFT tune estimator.

* ``seo_bilinear`` is the uniform-grid bilinear lookup with clamped
  indices (src/FFAG_MathTools.py:827-936) used to read r0(Ek, phi) and
  pr0(Ek, phi) from the SEO tables.
* ``rms_emittance`` follows compute_rms_emit (src/FFAG_Distribution.py).
* ``tune_from_xpx`` is the FFT tune estimator of the reference's space-charge
  tune benchmark (benchmarks/Benchmark_SC_kvTune.py:318-365): the complex
  signal x + i px / scale, zero-padded FFT, parabolic interpolation of the
  magnitude peak; ``fractional_tunes`` applies its fold (:368-378).
* ``load_position_dump`` reads the per-particle files a ``PositionDump``
  writes (``particle_{gid}_angle_{deg}_rank_0.npz``, 17 columns, the last
  one the time).
"""
import glob
import os
import re

import numpy as np


def seo_bilinear(ek_axis, fi_axis, table, ek, fi):
    """Bilinear value of ``table[i_ek, j_fi]`` at (ek, fi): floor index on the
    uniform axes, clamped to [0, n-2] (clamped points extrapolate from the
    edge cell).  Returns (values, out_of_range flags)."""
    ek = np.asarray(ek, dtype=np.float64)
    fi = np.asarray(fi, dtype=np.float64)
    x0, dx, nx = ek_axis[0], ek_axis[1] - ek_axis[0], len(ek_axis)
    y0, dy, ny = fi_axis[0], fi_axis[1] - fi_axis[0], len(fi_axis)
    i = np.floor((ek - x0) / dx).astype(np.int64)
    j = np.floor((fi - y0) / dy).astype(np.int64)
    out = (i < 0) | (i >= nx - 1) | (j < 0) | (j >= ny - 1)
    i = np.clip(i, 0, nx - 2)
    j = np.clip(j, 0, ny - 2)
    tx = (ek - (x0 + i * dx)) / dx
    ty = (fi - (y0 + j * dy)) / dy
    z = np.asarray(table)
    val = ((1 - tx) * (1 - ty) * z[i, j] + (1 - tx) * ty * z[i, j + 1] + tx * (1 - ty) * z[i + 1, j]
           + tx * ty * z[i + 1, j + 1])
    return val, out


def closed_orbit_offsets(rows, seo):
    """(x, x', z, z') of dumped rows (r, pr, z, pz, fi, Ek, ...) about the
    closed orbit of each particle's own energy: x = r - r0(Ek, phi mod 2pi),
    x' = pr - pr0(Ek, phi mod 2pi).  ``seo`` carries ``Ek_axis``,
    ``fi_axis``, ``r_matrix`` and ``pr_matrix``."""
    rows = np.asarray(rows, dtype=np.float64)
    fi = np.mod(rows[:, 4], 2 * np.pi)
    r0, _ = seo_bilinear(seo.Ek_axis, seo.fi_axis, seo.r_matrix, rows[:, 5], fi)
    pr0, _ = seo_bilinear(seo.Ek_axis, seo.fi_axis, seo.pr_matrix, rows[:, 5], fi)
    return rows[:, 0] - r0, rows[:, 1] - pr0, rows[:, 2], rows[:, 3]


def rms_emittance(x, xp):
    """sqrt(<dx^2><dxp^2> - <dx dxp>^2) about the means."""
    x = np.asarray(x, dtype=np.float64)
    xp = np.asarray(xp, dtype=np.float64)
    dx, dp = x - x.mean(), xp - xp.mean()
    return float(np.sqrt(np.mean(dx * dx) * np.mean(dp * dp) - np.mean(dx * dp) ** 2))


def _parabolic_peak(mag, k):
    y0, y1, y2 = mag[k - 1], mag[k], mag[k + 1]
    den = y0 - 2.0 * y1 + y2
    return float(k) if den == 0 else k + 0.5 * (y0 - y2) / den


def tune_from_xpx(x, px, pad_factor=8, ignore_dc=True, px_scale=None):
    """Raw tune in [0, 1) of a turn-by-turn signal: peak of |FFT| of
    x + i px / px_scale (px_scale = std(px) / std(x) unless given), with the
    signal's mean removed and the FFT zero-padded to pad_factor x N points,
    refined by a parabola through the peak and its neighbours."""
    x = np.asarray(x, dtype=np.float64)
    px = np.asarray(px, dtype=np.float64)
    n = x.size
    if n != px.size:
        raise ValueError("x and px must have same length")
    if n < 8:
        raise ValueError("too few TBT points for FFT tune analysis")
    if px_scale is None:
        sx, sp = x.std(), px.std()
        px_scale = sp / sx if sx > 0 and sp > 0 else 1.0
    zz = x + 1j * (px / px_scale)
    zz = zz - zz.mean()
    nfft = int(pad_factor * n)
    mag = np.abs(np.fft.fft(zz, n=nfft))
    if ignore_dc:
        mag[0] = 0.0
    k = int(np.argmax(mag))
    k = min(max(k, 1), nfft - 2)
    return (_parabolic_peak(mag, k) % nfft) / nfft


def fractional_tunes(vx_raw, vz_raw):
    """Fold of the raw FFT tunes into the plotted fractional tunes:
    q = 1 - v, and qz shifted down by one above 0.25."""
    qx = 1.0 - np.asarray(vx_raw, dtype=np.float64)
    qz = 1.0 - np.asarray(vz_raw, dtype=np.float64)
    qz = np.where(qz > 0.25, qz - 1.0, qz)
    return qx, qz


def load_position_dump(folder, angle_deg):
    """{global id: rows sorted by time} from a PositionDump folder."""
    out = {}
    pat = re.compile(r"particle_(\d+)_angle_" + re.escape(f"{angle_deg:.1f}") + r"_rank_\d+\.npz$")
    for path in glob.glob(os.path.join(folder, "particle_*_angle_*_rank_*.npz")):
        m = pat.search(os.path.basename(path))
        if not m:
            continue
        a = np.load(path)["particles"]
        out[int(m.group(1))] = a[np.argsort(a[:, -1], kind="stable")]
    return out


def particle_tunes(tbt, seo, turns=None):
    """Fractional (qx, qz) per particle from turn-by-turn rows about the
    closed orbit; particles with fewer than ``turns`` records are skipped
    (a lost particle), others are cut to the first ``turns`` records.
    Returns (ids, qx, qz)."""
    ids, qx, qz = [], [], []
    for gid in sorted(tbt):
        a = tbt[gid]
        if turns is not None:
            if a.shape[0] < turns:
                continue
            a = a[:turns]
        if a.shape[0] < 16:
            continue
        x, xp, z, zp = closed_orbit_offsets(a, seo)
        q = fractional_tunes(tune_from_xpx(x, xp), tune_from_xpx(z, zp))
        ids.append(gid)
        qx.append(float(q[0]))
        qz.append(float(q[1]))
    return np.asarray(ids, dtype=np.int64), np.asarray(qx), np.asarray(qz)
