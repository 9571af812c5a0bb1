"""Injected-bunch generation for the tracking entry (host side).

Same names, arguments and conventions as src/FFAG_Distribution.py:

* ``TwissContainer`` and the phase-space distributions ``KVDist1D``,
  ``WaterBagDist1D``, ``KVDist2D``, ``WaterBagDist2D``, ``GaussDist1D``,
  ``GaussDist2D``, ``HollowWaterBagDist1D``, ``IndependentHollowWaterBagDist2D``,
  ``CustomDist3D`` and ``create_distribution`` (:16-577, 1074-1150).  The
  input emittance of each is the rms emittance; the KV / water-bag boundary
  is at 2 / 4 (1D) and 4 / 6 (2D) times it.  Besides the reference's
  one-particle ``getCoordinates()``, each has ``sample(n)`` that draws n
  particles at once (numpy), which the bunch generator uses;
* ``LoadSEOParams``, ``LoadSEOBetaFunc``, ``GetInitRAndFiForInjection`` (for
  an injection azimuth, an injection centre line given as a polyline file,
  or an explicit {fi, r, pr} point), ``GenerateInitDistributionDtInjNew``
  with the RF-bucket-matched longitudinal type ``match`` (``FFAG_bucket``)
  and ``GenerateBunches`` with painting (:1152-1741).

Random numbers come from a numpy Generator (``set_seed`` makes a run
repeatable; the reference draws from the unseeded ``random`` module and, for
``match``, from fresh unseeded numpy generators).
"""
from __future__ import annotations

import math
import os
import re

import numpy as np

from .FFAG_Bunch import FFAG_ManageBunchAttribute
from .FFAG_bucket import FFAG_bucket

_rng = np.random.default_rng()


def set_seed(seed):
    """Reseed the generator all distributions draw from."""
    global _rng
    _rng = np.random.default_rng(seed)


class TwissContainer:
    """alpha, beta, gamma = (1 + alpha^2) / beta and the emittance; maps
    normalised (u_n, up_n) to u = sqrt(beta eps) u_n,
    up = sqrt(eps / beta) (up_n - alpha u_n)."""

    def __init__(self, alpha, beta, emittance):
        self.alpha, self.beta = alpha, beta
        self.setEmittance(emittance)

    def setEmittance(self, emittance):
        self.emittance = emittance
        self.gamma = (1.0 + self.alpha ** 2) / self.beta
        self.u_max = math.sqrt(self.beta * emittance)
        self.up_coeff = math.sqrt(emittance / self.beta)
        self.up_max = math.sqrt(self.gamma * emittance)

    def getNormalizedH(self, u, up):
        return (u ** 2 + (self.alpha * u + self.beta * up) ** 2) / (self.beta * self.emittance)

    def getU_Max(self):
        return self.u_max

    def getUP_Max(self):
        return self.up_max

    def getU_UP(self, u_norm, up_norm):
        return self.u_max * u_norm, self.up_coeff * (up_norm - self.alpha * u_norm)

    def getAlphaBetaGammaEmitt(self):
        return self.alpha, self.beta, self.gamma, self.emittance

    def getAlphaBetaEmitt(self):
        return self.alpha, self.beta, self.emittance


class _Dist:
    def getCoordinates(self):
        return tuple(float(v[0]) for v in self.sample(1))


class KVDist1D(_Dist):
    """Points on the ellipse of 2x the rms emittance."""

    def __init__(self, twiss=None):
        a, b, e = (twiss or TwissContainer(0.0, 1.0, 1.0)).getAlphaBetaEmitt()
        self.twiss = TwissContainer(a, b, 2 * e)
        self.emit_times = 2.0

    def sample(self, n):
        x = np.sin(2 * np.pi * (_rng.random(n) - 0.5))
        xp = _rng.choice((-1.0, 1.0), n) * np.sqrt(1.0 - x ** 2)
        return self.twiss.getU_UP(x, xp)

    def getTwissContainers(self):
        return (self.twiss,)


class WaterBagDist1D(_Dist):
    """Uniform inside the ellipse of 4x the rms emittance."""

    def __init__(self, twiss=None):
        a, b, e = (twiss or TwissContainer(0.0, 1.0, 1.0)).getAlphaBetaEmitt()
        self.kv_dist = KVDist1D(TwissContainer(a, b, 2 * e))
        self.emit_times = 2.0

    def sample(self, n):
        u, up = self.kv_dist.sample(n)
        g = np.sqrt(_rng.random(n))
        return g * u, g * up

    def getTwissContainers(self):
        return self.kv_dist.getTwissContainers()


class KVDist2D(_Dist):
    """Uniform on the 4D hyper-ellipsoid surface of 4x the rms emittances."""

    def __init__(self, twissX=None, twissY=None):
        ax, bx, ex = (twissX or TwissContainer(0.0, 1.0, 1.0)).getAlphaBetaEmitt()
        ay, by, ey = (twissY or TwissContainer(0.0, 1.0, 1.0)).getAlphaBetaEmitt()
        self.twissX = TwissContainer(ax, bx, 4 * ex)
        self.twissY = TwissContainer(ay, by, 4 * ey)
        self.emit_times = 4.0

    def sample(self, n):
        phi = 2 * np.pi * (_rng.random(n) - 0.5)
        rho = np.sqrt(_rng.random(n))
        p0 = np.sqrt(np.abs(1.0 - rho ** 2))
        psi = 2 * np.pi * (_rng.random(n) - 0.5)
        x, xp = self.twissX.getU_UP(rho * np.cos(phi), p0 * np.cos(psi))
        y, yp = self.twissY.getU_UP(rho * np.sin(phi), p0 * np.sin(psi))
        return x, xp, y, yp

    def getTwissContainers(self):
        return self.twissX, self.twissY


class WaterBagDist2D(_Dist):
    """Uniform inside the 4D hyper-ellipsoid of 6x the rms emittances."""

    def __init__(self, twissX=None, twissY=None):
        ax, bx, ex = (twissX or TwissContainer(0.0, 1.0, 1.0)).getAlphaBetaEmitt()
        ay, by, ey = (twissY or TwissContainer(0.0, 1.0, 1.0)).getAlphaBetaEmitt()
        self.kv_dist = KVDist2D(TwissContainer(ax, bx, 6.0 * ex / 4.0), TwissContainer(ay, by, 6.0 * ey / 4.0))
        self.emit_times = 6.0

    def sample(self, n):
        x, xp, y, yp = self.kv_dist.sample(n)
        g = np.sqrt(np.sqrt(_rng.random(n)))
        return g * x, g * xp, g * y, g * yp

    def getTwissContainers(self):
        return self.kv_dist.getTwissContainers()


class GaussDist1D(_Dist):
    """Gaussian of the rms emittance; cut_off > 0 rejects points beyond
    cut_off sigma in the normalised phase plane."""

    def __init__(self, twiss=None, cut_off=-1.0):
        self.twiss = twiss or TwissContainer(0.0, 1.0, 1.0)
        self.cut_off = cut_off
        self.cut_off2 = cut_off * cut_off
        self.emit_times = 1.0

    def sample(self, n):
        x = _rng.standard_normal(n)
        xp = _rng.standard_normal(n)
        if self.cut_off > 0.0:
            bad = x ** 2 + xp ** 2 > self.cut_off2
            while bad.any():
                k = int(bad.sum())
                x[bad] = _rng.standard_normal(k)
                xp[bad] = _rng.standard_normal(k)
                bad = x ** 2 + xp ** 2 > self.cut_off2
        return self.twiss.getU_UP(x, xp)

    def getTwissContainers(self):
        return (self.twiss,)


class GaussDist2D(_Dist):
    def __init__(self, twissX=None, twissY=None, cut_off=-1.0):
        self.twissX = twissX or TwissContainer(0.0, 1.0, 1.0)
        self.twissY = twissY or TwissContainer(0.0, 1.0, 1.0)
        self.gaussX = GaussDist1D(self.twissX, cut_off)
        self.gaussY = GaussDist1D(self.twissY, cut_off)
        self.cut_off = cut_off
        self.emit_times = 1.0

    def sample(self, n):
        x, xp = self.gaussX.sample(n)
        y, yp = self.gaussY.sample(n)
        return x, xp, y, yp

    def getTwissContainers(self):
        return self.twissX, self.twissY


class HollowWaterBagDist1D(_Dist):
    """Points of the ellipse of the given emittance scaled by g uniform in
    [0.4, 1]."""

    def __init__(self, twiss=None):
        a, b, e = (twiss or TwissContainer(0.0, 1.0, 1.0)).getAlphaBetaEmitt()
        self.kv_dist = KVDist1D(TwissContainer(a, b, e / 2))

    def sample(self, n):
        u, up = self.kv_dist.sample(n)
        g = _rng.uniform(0.4, 1.0, n)
        return g * u, g * up

    def getTwissContainers(self):
        return self.kv_dist.getTwissContainers()


class IndependentHollowWaterBagDist2D(_Dist):
    """Independent hollow 1D planes, kept when 0.3 <= (x/x_max)^2 +
    (y/y_max)^2 <= 1."""

    def __init__(self, twissX=None, twissY=None):
        self.hollow_x = HollowWaterBagDist1D(twissX)
        self.hollow_y = HollowWaterBagDist1D(twissY)

    def sample(self, n):
        ux = self.hollow_x.kv_dist.twiss.getU_Max()
        uy = self.hollow_y.kv_dist.twiss.getU_Max()
        out = [np.empty(0)] * 4
        while out[0].size < n:
            x, xp = self.hollow_x.sample(n)
            y, yp = self.hollow_y.sample(n)
            q = x ** 2 / ux ** 2 + y ** 2 / uy ** 2
            ok = (q >= 0.3) & (q <= 1.0)
            out = [np.concatenate((o, v[ok])) for o, v in zip(out, (x, xp, y, yp))]
        return tuple(o[:n] for o in out)

    def getTwissContainers(self):
        return self.hollow_x.getTwissContainers(), self.hollow_y.getTwissContainers()


class CustomDist3D(_Dist):
    """A 2D transverse and a 1D longitudinal distribution together."""

    def __init__(self, dist_transverse, dist_longitudinal):
        self.dist_transverse = dist_transverse
        self.dist_longitudinal = dist_longitudinal

    def sample(self, n):
        return tuple(self.dist_transverse.sample(n)) + tuple(self.dist_longitudinal.sample(n))

    def getTwissContainers(self):
        return tuple(self.dist_transverse.getTwissContainers()) + (self.dist_longitudinal.getTwissContainers(),)


def create_distribution(dist_type, twiss_transverse=None, twiss_longitudinal=None):
    """2D transverse distribution when ``twiss_transverse`` is given, else
    the 1D longitudinal one ('match' and 'UniformGauss' draw a Gaussian that
    the generator then replaces)."""
    two = twiss_transverse is not None
    if dist_type == "gauss":
        return GaussDist2D(*twiss_transverse) if two else GaussDist1D(twiss_longitudinal)
    if dist_type == "kv":
        return KVDist2D(*twiss_transverse) if two else KVDist1D(twiss_longitudinal)
    if dist_type == "waterbag":
        return WaterBagDist2D(*twiss_transverse) if two else WaterBagDist1D(twiss_longitudinal)
    if dist_type == "hollow_waterbag":
        return IndependentHollowWaterBagDist2D(*twiss_transverse) if two else HollowWaterBagDist1D(twiss_longitudinal)
    if dist_type in ("match", "UniformGauss"):
        return GaussDist1D(twiss_longitudinal)
    raise ValueError(f"Unknown distribution type: {dist_type}")


# ---------------------------------------------------------------------------
# closed-orbit and Twiss tables
# ---------------------------------------------------------------------------
def _bracket2(X, xi):
    """Indices of the two nodes bracketing xi (searchsorted 'right'; the
    first / last pair outside the axis)."""
    right = np.searchsorted(X, xi, side="right")
    return np.clip(right - 1, 0, len(X) - 2)


def _linear(X, Y, xi):
    i = _bracket2(X, xi)
    x0, x1, y0, y1 = X[i], X[i + 1], Y[i], Y[i + 1]
    return y0 + (xi - x0) / (x1 - x0) * (y1 - y0)


def _stencil4(X, xi):
    """First index of the 4-node stencil around xi: (left - 1 .. left + 2)
    with left from searchsorted 'right', moved inside the axis at the ends."""
    left = np.searchsorted(X, xi, side="right") - 1
    return np.clip(left - 1, 0, len(X) - 4)


def _lagrange_weights(xs, xi):
    w = np.ones((4,) + np.shape(xi))
    for a in range(4):
        for b in range(4):
            if a != b:
                w[a] = w[a] * (xi - xs[b]) / (xs[a] - xs[b])
    return w


def lagrange_2d(X, Y, Z, Xi, Yi):
    """Cubic Lagrange interpolation of Z[X, Y] on 4 x 4 nodes around each
    (Xi, Yi) (Lagrange_interp_2D_vect, src/FFAG_MathTools.py:274-314)."""
    if len(X) < 4 or len(Y) < 4:
        raise ValueError("cubic Lagrange interpolation needs at least 4 nodes per axis")
    Xi = np.atleast_1d(np.asarray(Xi, dtype=np.float64))
    Yi = np.atleast_1d(np.asarray(Yi, dtype=np.float64))
    ix, iy = _stencil4(X, Xi), _stencil4(Y, Yi)
    xs = np.stack([X[ix + a] for a in range(4)])
    ys = np.stack([Y[iy + b] for b in range(4)])
    wx, wy = _lagrange_weights(xs, Xi), _lagrange_weights(ys, Yi)
    out = np.zeros_like(Xi)
    for a in range(4):
        for b in range(4):
            out += wx[a] * wy[b] * Z[ix + a, iy + b]
    return out


def LoadSEOParams(SEOFileName, Ek_c):
    """(Ek_c, r0, pr0, T_rev) at Ek_c by linear interpolation over the
    SEO_ini.txt rows (columns Ek 1, f_rev 4, r0 6, pr0 7), the folder two
    levels up (the Bmap folder), and the SEO_ini table."""
    path = os.path.dirname(os.path.dirname(SEOFileName))
    data = np.loadtxt(SEOFileName, skiprows=2, ndmin=2)
    ek, r0, pr0, t0 = data[:, 1], data[:, 6], data[:, 7], 1 / data[:, 4]
    return np.array([Ek_c, _linear(ek, r0, Ek_c), _linear(ek, pr0, Ek_c), _linear(ek, t0, Ek_c)]), path, data


def LoadSEOBetaFunc(SEOPATHName, Ek_c, fi_c):
    """(beta_z, beta_r, alpha_z, alpha_r) at (Ek_c, fi_c) from
    BetaFuncZ/BetaFuncR/AlphaFuncZ/AlphaFuncR.txt (first column Ek, first
    row the azimuth axis).  As in the reference, fi_c is taken modulo
    deg2rad of the axis span before the lookup."""
    out = []
    for name in ("BetaFuncZ.txt", "BetaFuncR.txt", "AlphaFuncZ.txt", "AlphaFuncR.txt"):
        d = np.loadtxt(os.path.join(SEOPATHName, name), skiprows=1)
        ek_axis, fi_axis, m = d[1:, 0], d[0, 1:], d[1:, 1:]
        fi_m = np.mod(np.array([fi_c]), np.deg2rad(fi_axis[-1] - fi_axis[0]))
        out.append(float(lagrange_2d(ek_axis, fi_axis, m, np.array([Ek_c]), fi_m)[0]))
    return tuple(out)


def _polyline_segments(pts):
    """(n, 2) points -> (n - 1, 4) segments (x1, y1, x2, y2)
    (CentralLine2Segments, src/FFAG_Utils.py:278-286)."""
    pts = np.asarray(pts, dtype=np.float64)
    return np.column_stack((pts[:-1, 0], pts[:-1, 1], pts[1:, 0], pts[1:, 1]))


def _segment_crossings(A, B):
    """Proper crossings of every segment of A with every segment of B
    (FindIntersectionVect, src/FFAG_Utils.py:32-152): the bounding boxes
    overlap strictly and each segment's end points lie strictly on either
    side of the other's line.  Returns the (len(A), len(B)) crossing flags
    and the signed distances of A's two end points from B's lines."""
    a1x, a1y, a2x, a2y = (A[:, k][:, None] for k in range(4))
    b1x, b1y, b2x, b2y = (B[:, k][None, :] for k in range(4))
    apart = ((np.minimum(a1x, a2x) >= np.maximum(b1x, b2x)) | (np.maximum(a1x, a2x) <= np.minimum(b1x, b2x)) |
             (np.minimum(a1y, a2y) >= np.maximum(b1y, b2y)) | (np.maximum(a1y, a2y) <= np.minimum(b1y, b2y)))
    ca1 = (b2x - b1x) * (a1y - b1y) - (b2y - b1y) * (a1x - b1x)
    ca2 = (b2x - b1x) * (a2y - b1y) - (b2y - b1y) * (a2x - b1x)
    cb1 = (a2x - a1x) * (b1y - a1y) - (a2y - a1y) * (b1x - a1x)
    cb2 = (a2x - a1x) * (b2y - a1y) - (a2y - a1y) * (b2x - a1x)
    lb = np.hypot(b2x - b1x, b2y - b1y)
    with np.errstate(divide="ignore", invalid="ignore"):
        d1 = np.where(lb != 0, ca1 / lb, np.nan)
        d2 = np.where(lb != 0, ca2 / lb, np.nan)
    return ~apart & (ca1 * ca2 < 0) & (cb1 * cb2 < 0), d1, d2


def GetInitRAndFiForInjection(SEO_R, SEO_Pr, InjectCL, Ek):
    """(theta_rad, r, pr) of the injection point on the closed orbit of
    energy Ek.  InjectCL is either the injection azimuth in degrees
    (cubic Lagrange lookup of r and pr) or an (n, 2) polyline (x, y) [m] of
    the injection centre line (src/FFAG_Distribution.py:1347-1414): the orbit
    at Ek (SEO_R / SEO_Pr rows interpolated linearly in Ek) is drawn as a
    polyline through its azimuth nodes, closed back to its second node as the
    reference closes it, and the first orbit segment (then the first line
    segment) that the line crosses gives the point, placed on the orbit
    segment in the ratio of its end points' distances from the line; theta
    is its atan2 azimuth in (-pi, pi], r its radius and pr the orbit's pr at
    the segment's first node.  (The reference's own code for this case ends
    in an IndexError when it indexes the scalar azimuth, line 1412; this is
    the computation that code sets up.)"""
    ek_axis, fi_axis = SEO_R[1:, 0], SEO_R[0, 1:]
    if isinstance(InjectCL, np.ndarray):
        i = int(np.clip(np.searchsorted(ek_axis, Ek, side="right") - 1, 0, len(ek_axis) - 2))
        x0, x1 = ek_axis[i], ek_axis[i + 1]
        w0, w1 = (Ek - x1) / (x0 - x1), (Ek - x0) / (x1 - x0)
        r_row = w0 * SEO_R[1 + i, 1:] + w1 * SEO_R[2 + i, 1:]
        pr_row = w0 * SEO_Pr[1 + i, 1:] + w1 * SEO_Pr[2 + i, 1:]
        xy = np.column_stack((r_row * np.cos(fi_axis), r_row * np.sin(fi_axis)))
        xy = np.vstack((xy, xy[1]))
        seo_seg = _polyline_segments(xy)
        cl = np.atleast_2d(np.asarray(InjectCL, dtype=np.float64))[:, :2]
        hit, d1, d2 = _segment_crossings(seo_seg, _polyline_segments(cl))
        ia, ib = np.nonzero(hit)
        if ia.size == 0:
            raise ValueError(f"the injection centre line does not cross the closed orbit of Ek = {Ek} MeV")
        a, b = ia[0], ib[0]
        pre, post = abs(d1[a, b]), abs(d2[a, b])
        w = pre / (pre + post)
        p = (1.0 - w) * seo_seg[a, 0:2] + w * seo_seg[a, 2:4]
        return float(np.arctan2(p[1], p[0])), float(np.hypot(p[0], p[1])), float(pr_row[a])
    theta = np.array([np.deg2rad(InjectCL)])
    r = lagrange_2d(ek_axis, fi_axis, SEO_R[1:, 1:], np.array([Ek]), theta)[0]
    pr = lagrange_2d(ek_axis, fi_axis, SEO_Pr[1:, 1:], np.array([Ek]), theta)[0]
    return theta[0], r, pr


def GenerateInitDistributionDtInjNew(SEOFileName, InjectionPosition, Ek_c, Emit, num_points, InjTimeNanoSec,
                                     LongitudeAlpha, LongitudeBeta, LongitudeT, acc_U0=100.0e3, acc_phi0=0.0,
                                     TwissParams=None, dist_types=("gauss", "gauss"), dx_xp_y_yp=(0.0, 0.0, 0.0, 0.0),
                                     Bunch_ID=0.0, bucket_obj=None):
    """One sub-bunch as a (num_points, 16) matrix in the (r, pr, z, pz, fi,
    Ek, inj_t, ...) layout, centred on the closed orbit at the injection
    point plus the painting offsets.  Emit = (R, Z rms emittances in
    pi mm mrad, longitudinal emittance); the longitudinal coordinates are
    (inj_t [ns], dEk [MeV])."""
    eps_r, eps_z, eps_l = Emit[0] / 1e6, Emit[1] / 1e6, Emit[2]
    seo_ini, bmap_path, _ = LoadSEOParams(os.path.join(SEOFileName, "SEO_ini.txt"), Ek_c)
    seo_r = np.loadtxt(os.path.join(SEOFileName, "SEO_r.txt"), skiprows=1)
    seo_pr = np.loadtxt(os.path.join(SEOFileName, "SEO_pr.txt"), skiprows=1)
    Ek_c = seo_ini[0]
    if isinstance(InjectionPosition, str):
        cl = np.loadtxt(InjectionPosition, skiprows=1, ndmin=2)
        fi_c, r_c, pr_c = GetInitRAndFiForInjection(seo_r, seo_pr, cl, Ek_c)
    elif isinstance(InjectionPosition, dict):
        fi_c, r_c, pr_c = InjectionPosition["fi"], InjectionPosition["r"], InjectionPosition["pr"]
    elif isinstance(InjectionPosition, (int, float)):
        fi_c, r_c, pr_c = GetInitRAndFiForInjection(seo_r, seo_pr, float(InjectionPosition), Ek_c)
    else:
        raise ValueError("Unsupported type for InjectionPosition")
    if TwissParams is not None:
        bz, br, az, ar = TwissParams
    else:
        bz, br, az, ar = LoadSEOBetaFunc(SEOFileName, Ek_c, fi_c)
    t_type, l_type = dist_types
    if l_type == "kv":
        eps_l *= 2
    tw_r, tw_z = TwissContainer(ar, br, eps_r), TwissContainer(az, bz, eps_z)
    tw_l = TwissContainer(LongitudeAlpha, LongitudeBeta, eps_l)
    dist = CustomDist3D(create_distribution(t_type, twiss_transverse=(tw_r, tw_z)),
                        create_distribution(l_type, twiss_longitudinal=tw_l))
    r, pr, z, pz, t_inj, ek = (np.asarray(v, dtype=np.float64) for v in dist.sample(num_points))
    if l_type == "match":
        # inside the RF bucket, H above the level whose centre-line
        # trajectory spans LongitudeT ns (src/FFAG_Distribution.py:1275-1280)
        if bucket_obj is None:
            bucket_obj = FFAG_bucket(Bmapfoldname=bmap_path, Ek0=Ek_c, fi0=acc_phi0, U0=acc_U0, step_t=20e-9,
                                     step_N=18000)
        t_ns, ek_mev, _ = bucket_obj.scatter_filter_by_span(num_points, LongitudeT, show=False, rng=_rng)
        t_inj, ek = t_ns, ek_mev - Ek_c
    elif l_type == "UniformGauss":
        t_inj = _rng.uniform(-0.5 * LongitudeT, 0.5 * LongitudeT, num_points)
        ek = _rng.normal(0.0, math.sqrt(eps_l / LongitudeBeta), num_points)
    centre = (r_c + dx_xp_y_yp[0], pr_c + dx_xp_y_yp[1], dx_xp_y_yp[2], dx_xp_y_yp[3], Ek_c, InjTimeNanoSec)
    A = FFAG_ManageBunchAttribute().Attribute
    M = np.zeros((num_points, len(A)))
    M[:, A["r"]] = r + centre[0]
    M[:, A["vr"]] = pr + centre[1]
    M[:, A["z"]] = z + centre[2]
    M[:, A["vz"]] = pz + centre[3]
    M[:, A["fi"]] = fi_c
    M[:, A["Ek"]] = ek + Ek_c
    M[:, A["inj_t"]] = t_inj + InjTimeNanoSec
    M[:, A["Bunch_ID"]] = Bunch_ID
    M[:, A["Local_ID"]] = np.arange(num_points)
    M[:, A["Global_ID"]] = np.arange(num_points)
    return M, dist, centre, bucket_obj


def GenerateBunches(BunchPara):
    """All sub-bunches of the injection, stacked: one bunch, or with
    painting min(floor(max curve time / interval), MaxBunchNum) sub-bunches
    injected every PaintTimeInterval ns with the (x, px, y, py) offsets of the
    paint curve (columns t, x, px, y, py) interpolated at their time."""
    with open(BunchPara["EmapPATHName"]) as f:
        text = f.read()
    m_phi = re.search(r"acc_phi_paint = ([\d.eE+-]+)", text)
    m_v0 = re.search(r"acc_voltage_paint = ([\d.eE+-]+)", text)
    acc_phi0 = float(m_phi.group(1)) if m_phi else None
    acc_v0 = float(m_v0.group(1)) if m_v0 else None
    if BunchPara.get("PlotFlag"):
        raise NotImplementedError("bunch plotting (check.CheckBunch) is not supported")
    interval = BunchPara["PaintTimeInterval"]
    if BunchPara["PaintEnable"]:
        curve = np.loadtxt(BunchPara["PaintCurve"], skiprows=1, ndmin=2)
        n_sub = int(min(np.floor(np.max(curve[:, 0]) / interval), BunchPara["PaintMaxNum"]))
    else:
        curve, n_sub = None, 1
    t_len, d_ek = BunchPara["LongitudeT"], BunchPara["LongitudeDEk"]
    bunches = []
    for i in range(n_sub):
        t_i = i * interval
        if curve is not None:
            off = tuple(float(np.interp(t_i, curve[:, 0], curve[:, c])) for c in (1, 2, 3, 4))
        else:
            off = (0.0, 0.0, 0.0, 0.0)
        sub, _, _, _ = GenerateInitDistributionDtInjNew(
            BunchPara["SEO"], BunchPara["InjPosition"], BunchPara["InjectEk"],
            [BunchPara["TransverseREmit"], BunchPara["TransverseZEmit"], t_len * d_ek / 16],
            BunchPara["ParticleNum"], t_i, 0.0, t_len / d_ek, t_len, acc_U0=acc_v0, acc_phi0=acc_phi0,
            dist_types=(BunchPara["TransverseDistType"], BunchPara["LongitudeDistType"]), dx_xp_y_yp=off,
            Bunch_ID=i)
        bunches.append(sub)
    return np.vstack(bunches)
