"""Oracle (test infrastructure only): CPU restatement of the reference's
closed-orbit, tune and Twiss computations, used by tests to check the GPU
orbit tracker and the SEO finder.  Nothing in the product imports this
module.

Follows src/FFAG_track.py and src/FFAG_BetaFunc.py:
* ``rhs``        FunctionForSEOVect (:1641-1695): state (r, pr, z, pz),
                 theta the independent variable, order-0 field
                 (GetBfields, the C ``ora_get_bfields``) plus harmonics;
* ``rk4_solve``  rk4_step / rk4_solve (:33-40, :621-640) on the
                 linspace(t0, t1, int((t1 - t0) / h)) grid;
* ``closed_orbit``  SearchSEOUsingInitialEkVect + TrackBunchSEO
                 (:1167-1323): midpoint iteration over one sector, then a
                 finite-difference Newton iteration;
* ``tunes``      GetQrQz (:1325-1371) and GetFreq (:1425-1439);
* ``twiss``      CalcBetaFunc (src/FFAG_BetaFunc.py:26-190): one-sector
                 matrices from two test particles per plane, beta = a12 /
                 sin(mu), alpha = (a11 - a22) / (2 sin(mu)).

Pinned by tests/test_orbit_oracle.py: a uniform-flutter field of index k
has the closed forms r0^(k+1) = P / (q B0), Qr = sqrt(1 + k),
Qz = sqrt(-k), beta = r0 / Q, alpha = 0, perimeter 2 pi r0; on the bench
lattice the closed orbit matches the lattice module's independent (r, angle)
formulation and the tunes match FFT tunes of the oracle's time-stepped 6D
RK4 tracker; a planted sign slip in the Bfi term fails the latter.
"""
import math

import numpy as np

from . import oracle as O

Q = 1.60217662e-19
C = 2.99792458e8
E0_MEV = 938.2723


def ek_to_p(ek_mev):
    """Momentum [kg m/s] of a proton of kinetic energy ek (MeV)."""
    g = 1.0 + ek_mev / E0_MEV
    return g * math.sqrt(1.0 - 1.0 / (g * g)) * (E0_MEV * 1e6 * Q / C)


def ek_to_v(ek_mev):
    g = 1.0 + ek_mev / E0_MEV
    return C * math.sqrt(1.0 - 1.0 / (g * g))


def fields(mc, r, fi, z):
    """(Bz, Br, Bfi) of the order-0 field at fi mod the sector plus the
    harmonics at the raw fi."""
    L = O.lib()
    n = r.size
    M = np.zeros((n, 16))
    M[:, 0], M[:, 2], M[:, 4] = r, z, fi
    M[:, 7] = M[:, 8] = 1.0
    fim = np.mod(fi, 2 * np.pi / mc.nsectors)
    Bz, Br, Bfi = np.zeros(n), np.zeros(n), np.zeros(n)
    L.ora_get_bfields(O._p(M), O._i64(n), O._p(np.ascontiguousarray(fim)), O._p(Bz), O._p(Br), O._p(Bfi),
                      O._p(mc.s_poly), O._i(mc.s_poly.size), O._p(mc.sp_poly), O._i(mc.sp_poly.size), O._d(mc.B0),
                      O._d(mc.k_value), O._i(mc.nsectors), O._p(mc.convex), O._i(mc.convex.shape[0]))
    if mc.harm_enable:
        L.ora_bharmonics(O._p(M), O._i64(n), O._i(mc.harm_m.size), O._p(mc.harm_m), O._p(mc.harm_amp),
                         O._p(mc.harm_phase), O._p(mc.harm_rmin), O._p(mc.harm_rmax), O._p(Bz), O._p(Br), O._p(Bfi))
    return Bz, Br, Bfi


def rhs(mc, t, X, qp, bfi_sign=1.0):
    """d(r, pr, z, pz)/dtheta; ``bfi_sign`` = -1 plants a sign slip in the
    Bfi terms (a test's negative control)."""
    fi = np.full(X.shape[0], t % (np.pi * 2))
    r, pr, z, pz = X[:, 0], X[:, 1], X[:, 2], X[:, 3]
    Bz, Br, Bfi = fields(mc, r, fi, z)
    Bfi = bfi_sign * Bfi
    pf = np.sqrt(1 - pr ** 2 - pz ** 2)
    d = np.empty_like(X)
    d[:, 0] = r * pr / pf
    d[:, 1] = pf - qp * (r * Bz - r * pz / pf * Bfi)
    d[:, 2] = r * pz / pf
    d[:, 3] = qp * (r * Br - r * pr / pf * Bfi)
    return d


def rk4_solve(mc, X0, qp, t0, t1, h, record=False, **kw):
    """States on the linspace grid (trajectory if ``record``, else the end)."""
    steps = int((t1 - t0) / h)
    tp = np.linspace(t0, t1, steps)
    hh = tp[1] - tp[0]
    X = np.array(X0, dtype=np.float64, copy=True)
    qp = np.asarray(qp, dtype=np.float64)
    traj = [X.copy()] if record else None
    for i in range(len(tp) - 1):
        t = tp[i]
        k1 = hh * rhs(mc, t, X, qp, **kw)
        k2 = hh * rhs(mc, t + 0.5 * hh, X + 0.5 * k1, qp, **kw)
        k3 = hh * rhs(mc, t + 0.5 * hh, X + 0.5 * k2, qp, **kw)
        k4 = hh * rhs(mc, t + hh, X + k3, qp, **kw)
        X = X + (k1 + 2 * k2 + 2 * k3 + k4) / 6
        if record:
            traj.append(X.copy())
    return (tp, np.array(traj)) if record else X


def closed_orbit(mc, ek, r_guess, h=0.001, **kw):
    """(r0, pr0) at azimuth 0 of the closed orbit of energy ek."""
    qp = Q / ek_to_p(ek)
    sector = np.pi * 2 / mc.nsectors
    r0, pr0 = r_guess, 0.0
    for _ in range(200):
        e = rk4_solve(mc, [[r0, pr0, 0.0, 0.0]], [qp], 0.0, sector, h, **kw)[0]
        if abs(e[0] - r0) < 1e-5 and abs(e[1] - pr0) < 1e-5:
            break
        r0, pr0 = (r0 + e[0]) / 2, (pr0 + e[1]) / 2
    dr, dp = r0 * 1e-4, 1e-3
    for _ in range(25):
        X = np.array([[r0, pr0, 0, 0], [r0 + dr, pr0, 0, 0], [r0, pr0 + dp, 0, 0]], dtype=np.float64)
        E = rk4_solve(mc, X, [qp] * 3, 0.0, sector, h, **kw)
        (r0f, p0f), (r1f, p1f), (r2f, p2f) = E[0, :2], E[1, :2], E[2, :2]
        a11, a12 = (r1f - r0f) / dr, (r2f - r0f) / dp
        a21, a22 = (p1f - p0f) / dr, (p2f - p0f) / dp
        a11p, a22p = a11 - 1, a22 - 1
        det = a11p * a22p - a12 * a21
        if abs((r0f - r0) / r0) < 1e-8:
            break
        r0, pr0 = (r0 + a22p / det * (r0 - r0f) - a12 / det * (pr0 - p0f),
                   pr0 + a11p / det * (pr0 - p0f) - a21 / det * (r0 - r0f))
    return r0, pr0


def tunes(mc, r0, pr0, ek, h=0.001, **kw):
    """(Qr, Qz, f_rev, mean r, perimeter) from one-sector matrices."""
    qp = Q / ek_to_p(ek)
    sector = np.pi * 2 / mc.nsectors
    drr, dpp, dzz, dpz = 1e-6, 1e-6, 1e-7, 1e-7
    X = np.array([[r0, pr0, 0, 0], [r0 + drr, pr0, 0, 0], [r0, pr0 + dpp, 0, 0], [r0, pr0, dzz, 0],
                  [r0, pr0, 0, dpz]], dtype=np.float64)
    tp, tr = rk4_solve(mc, X, [qp] * 5, 0.0, sector, h, record=True, **kw)
    E = tr[-1]
    a11 = (E[1, 0] - E[0, 0]) / drr
    a22 = (E[2, 1] - E[0, 1]) / dpp
    b11 = (E[3, 2] - 0.0) / dzz
    b22 = (E[4, 3] - 0.0) / dpz
    qr = np.arccos((a11 + a22) / 2) / (2 * np.pi) * mc.nsectors
    qz = np.arccos((b11 + b22) / 2) / (2 * np.pi) * mc.nsectors
    r, pr = tr[:, 0, 0], tr[:, 0, 1]
    rp = pr / np.sqrt(1 - pr ** 2) * r
    dl = np.sqrt(r ** 2 + rp ** 2) * (tp[2] - tp[1])
    perim = np.sum(dl[:-1]) * mc.nsectors
    return qr, qz, ek_to_v(ek) / perim, float(np.mean(r)), perim


def _coef(s1, ps1, e1, pe1, s2, ps2, e2, pe2):
    A = np.array([[s1, ps1, 0, 0], [0, 0, s1, ps1], [s2, ps2, 0, 0], [0, 0, s2, ps2]])
    out = []
    for col in range(4):
        B = A.copy()
        B[:, col] = [e1, pe1, e2, pe2]
        out.append(np.linalg.det(B) / np.linalg.det(A))
    return out


def twiss(mc, r0, pr0, ek, fi0, h=0.001):
    """(beta_r, beta_z, alpha_r, alpha_z) at the azimuths fi0 (within the
    first sector) from a two-sector track of 5 particles."""
    qp = Q / ek_to_p(ek)
    sector = np.pi * 2 / mc.nsectors
    d = 1e-7
    X = np.array([[r0, pr0, 0, 0], [r0 + d, pr0, 0, 0], [r0, pr0 + d, 0, 0], [r0, pr0, d, 0], [r0, pr0, 0, d]],
                 dtype=np.float64)
    tp, tr = rk4_solve(mc, X, [qp] * 5, 0.0, 2 * sector, h, record=True)

    def at(t):
        return np.stack([np.stack([np.interp(t, tp, tr[:, p, c]) for c in range(4)]) for p in range(5)])
    out = []
    for f in np.atleast_1d(fi0):
        s, e = at(f), at(f + sector)
        ds, de = s[1:] - s[0], e[1:] - e[0]
        a = _coef(ds[0, 0], ds[0, 1], de[0, 0], de[0, 1], ds[1, 0], ds[1, 1], de[1, 0], de[1, 1])
        b = _coef(ds[2, 2], ds[2, 3], de[2, 2], de[2, 3], ds[3, 2], ds[3, 3], de[3, 2], de[3, 3])
        row = []
        for m11, m12, m21, m22 in (a, b):
            c = (m11 + m22) / 2
            sn = math.sqrt(1 - c * c)
            if m12 * sn < 0:
                sn = -sn
            row.append((m12 / sn, (m11 - m22) / sn / 2))
        out.append((row[0][0], row[1][0], row[0][1], row[1][1]))
    return np.array(out)
