"""Synthetic lattice inputs for the tracking hot path.

The reference ships no field maps, closed-orbit (SEO) tables or a usable RF
frequency curve (its ``Bmap/`` folder is empty and its ``Emap/*.txt`` files use
an older 3-column layout), so this module generates them, in the reference's
file formats, for a scaling FFAG:

* ``config_Bmap.json`` -- schema of the reference's root ``config_Bmap.json``
  (read by ``src/FFAG_Field.py:636-706``).
* ``Bz/Br/Bfi_order_<n>.npz`` Taylor-coefficient stacks with keys
  ``Bz_Taylor_coeff`` etc. (loaded at ``src/FFAG_Field.py:691-706``).  Their
  content follows the reference generator's definition
  (``src/FFAG_HighOrder.py:936-998``):
  ``Bz_n = (-1)^n/(2n)! Lap^n B0z``, ``Br_n = (-1)^n/(2n+1)! d/dr Lap^n B0z``,
  ``Bfi_n = (-1)^n/(2n+1)! (1/r) d/dtheta Lap^n B0z``, evaluated here in closed
  form for a zero spiral angle, where ``B0z = B0 r^k F(theta)`` separates and
  ``Lap^n (r^k F) = r^(k-2n) G_n(theta)`` with ``G_0 = F`` and
  ``G_(m+1) = (k-2m)^2 G_m + G_m''``.
* SEO tables (``SEO_ini.txt``, ``SEO_r.txt``, ``SEO_pr.txt``) with the column
  layout ``LoadSEOParams`` reads (``src/FFAG_Distribution.py:1417-1427``) and
  ``examples/main_track.py:241-245``: closed orbits found here by a Newton
  search on a theta-integrated mid-plane orbit.
* A 4-column RF frequency-curve file (time, f_RF, V0, Ek) with the header keys
  written by ``examples/main_Emap.py:754-787`` and parsed by
  ``src/FFAG_Field.py:316-376``.
* A seeded particle distribution in the reference's 16-column
  ``(r, pr, z, pz, fi, Ek, inj_t, ...)`` layout (``src/FFAG_Bunch.py:23-62``).
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

Q_E = 1.60217662e-19          # src/FFAG_ParasAndConversion.py:13
C_LIGHT = 2.99792458e8        # src/FFAG_ParasAndConversion.py:14
E0_MEV = 938.2723             # src/FFAG_ParasAndConversion.py:15
E0_J = E0_MEV * 1e6 * Q_E


def ek_to_p(ek_mev):
    """Momentum [kg m/s] of a proton of kinetic energy ``ek_mev``
    (``FFAG_ConversionTools.Ek2P``, src/FFAG_ParasAndConversion.py:88-92)."""
    ek_j = np.asarray(ek_mev) * 1e6 * Q_E
    return np.sqrt((ek_j + E0_J) ** 2 - E0_J ** 2) / C_LIGHT


def ek_to_v(ek_mev):
    """Speed [m/s] (``Ek2v``, src/FFAG_ParasAndConversion.py:95-101)."""
    gamma = (np.asarray(ek_mev) * 1e6 * Q_E + E0_J) / E0_J
    return np.sqrt(1.0 - (1.0 / gamma) ** 2) * C_LIGHT


# ---------------------------------------------------------------------------
# Magnet description
# ---------------------------------------------------------------------------
@dataclass
class ScalingFFAG:
    """A scaling FFAG sector lattice: ``Bz = B0 r^k F(theta)`` (reference
    convention ``B0 * r**k``, src/FFAG_Field.py:1381-1386)."""
    nsector: int = 12
    r0: float = 8.5
    k: float = 7.6
    ek_ref_mev: float = 300.0
    rmin: float = 8.0
    rmax: float = 9.3
    rstep: float = 0.001
    theta_step_deg: float = 0.005
    intervals: tuple = (7.5 / 30, 9.375 / 30, 2.8125 / 30, 2.8125 / 30, 7.5 / 30)
    polarities: tuple = (0.0, 1.0, 0.0, -0.728, 0.0)
    fringes: tuple = (0.0, 0.20, 0.0, 0.35, 0.0)
    spiral_deg: float = 0.0
    B0: float | None = None
    folder: str = "."

    # -- geometry shared with FFAG_Bfield_analytical (src/FFAG_Field.py:655-678)
    @property
    def theta_step_rad(self):
        return math.radians(self.theta_step_deg)

    @property
    def n_fi(self):
        return int((2 * np.pi / self.nsector) / self.theta_step_rad) + 1

    @property
    def n_r(self):
        return int((self.rmax - self.rmin) / self.rstep) + 1

    def r_axis(self):
        return np.linspace(self.rmin, self.rmax, self.n_r)

    def fi_axis(self):
        return np.linspace(0, 2 * np.pi / self.nsector, self.n_fi)

    def convex_params(self):
        """Sigmoid segment table ``(x0, x1, W0, |polarity|, sign)`` in radians,
        one row per non-zero polarity, with the same meaning as
        ``BmapParams_to_ConvexParams`` (src/FFAG_Field.py:732-767)."""
        rows = []
        sector_deg = 360.0 / self.nsector
        a = 0.0
        for frac, pol, fr in zip(self.intervals, self.polarities, self.fringes):
            b = a + frac * sector_deg
            if pol != 0:
                rows.append((math.radians(a), math.radians(b),
                             math.radians((b - a) * fr), abs(pol),
                             1.0 if pol > 0 else -1.0))
            a = b
        return np.array(rows, dtype=np.float64)

    def flutter(self, theta):
        """F(theta) of the order-0 field (same sum of sigmoid pairs as
        ``GetBfields_Njit``, src/FFAG_Field.py:1393-1418)."""
        theta = np.asarray(theta, dtype=np.float64)
        out = np.zeros_like(theta)
        for x0, x1, w0, y01, y02 in self.convex_params():
            ks = 16.0 / w0
            s1 = 1.0 / (1.0 + np.exp(-ks * (theta - x0)))
            s2 = 1.0 / (1.0 + np.exp(ks * (theta - x1)))
            out += y02 * abs(y01) * s1 * s2
        return out

    def flutter_derivs(self, theta, m_max):
        """Exact derivatives F^(m)(theta), m = 0..m_max, as a (m_max+1, len)
        array.  d^m sigma(a t)/dt^m = a^m P_m(sigma) with P_0(s) = s and
        P_(m+1)(s) = P_m'(s) s (1-s); the product of the two sigmoids of a
        segment is differentiated with the Leibniz rule."""
        theta = np.asarray(theta, dtype=np.float64)
        polys = [np.polynomial.Polynomial([0.0, 1.0])]
        s1m = np.polynomial.Polynomial([0.0, 1.0, -1.0])  # s(1-s)
        for _ in range(m_max):
            polys.append(polys[-1].deriv() * s1m)
        out = np.zeros((m_max + 1, theta.size))
        for x0, x1, w0, y01, y02 in self.convex_params():
            a = 16.0 / w0
            u = 1.0 / (1.0 + np.exp(-a * (theta - x0)))
            v = 1.0 / (1.0 + np.exp(a * (theta - x1)))
            du = [a ** i * polys[i](u) for i in range(m_max + 1)]
            dv = [(-a) ** i * polys[i](v) for i in range(m_max + 1)]
            amp = y02 * abs(y01)
            for m in range(m_max + 1):
                acc = np.zeros_like(theta)
                for i in range(m + 1):
                    acc += math.comb(m, i) * du[i] * dv[m - i]
                out[m] += amp * acc
        return out

    def mean_flutter(self):
        th = np.linspace(0, 2 * np.pi / self.nsector, 20001)
        return float(np.mean(self.flutter(th)))

    def bz_midplane(self, r, theta):
        return self.B0 * np.asarray(r) ** self.k * self.flutter(np.mod(theta, 2 * np.pi / self.nsector))

    def to_config(self):
        """config_Bmap.json content in the reference schema (root
        ``config_Bmap.json``)."""
        return {
            "Bmap": {
                "B0_T": self.B0,
                "Bmean": None,
                "NSector": self.nsector,
                "R0_m": self.r0,
                "SpiralAngle_deg": self.spiral_deg,
                "Type": "SCALE",
                "coefficients_tupel": [self.B0, self.k],
                "fringe_width_1": list(self.fringes),
                "interval_1": list(self.intervals),
                "k_value": self.k,
                "positive_or_negative_1": list(self.polarities),
                "r_axis": None,
                "rmin_max_step_m": [self.rmin, self.rmax, self.rstep],
                "theta_step_rad": self.theta_step_rad,
            },
            "BzInfo": {"IsoCoef": None, "IsoError": None},
            "config": {"date": 0.0, "expand_order": 0, "folder": self.folder},
            "machine": {"energy_inj": self.ek_ref_mev},
        }


def solve_B0(lat: ScalingFFAG, iters: int = 3) -> float:
    """Choose B0 so that the ``ek_ref_mev`` closed orbit has its mean radius at
    ``r0``: start from the hard-edge estimate ``Brho = r0 * B0 r0^k <F>`` and
    correct with the scaling law ``r_co ~ B0^(-1/(k+1))``."""
    brho = ek_to_p(lat.ek_ref_mev) / Q_E
    lat.B0 = brho / (lat.r0 ** (lat.k + 1) * lat.mean_flutter())
    for _ in range(iters):
        co = closed_orbit(lat, lat.ek_ref_mev, n_theta=4000)
        rmean = float(np.mean(co["r"]))
        lat.B0 *= (rmean / lat.r0) ** (lat.k + 1)
    return lat.B0


# ---------------------------------------------------------------------------
# Taylor maps
# ---------------------------------------------------------------------------
def _g_coeffs(k, n_max):
    """Coefficients c[n][i] with G_n = sum_i c[n][i] F^(2i)."""
    coeffs = [np.array([1.0])]
    for m in range(n_max):
        c = coeffs[-1]
        nxt = np.zeros(c.size + 1)
        nxt[: c.size] += (k - 2 * m) ** 2 * c
        nxt[1:] += c
        coeffs.append(nxt)
    return coeffs


def taylor_maps(lat: ScalingFFAG, max_order: int, dtype=np.float64):
    """Taylor stacks ``(max_order+1, NR, NPHI)`` for Bz, Br, Bfi (zero-spiral
    closed form of the reference generator's definition)."""
    if lat.spiral_deg != 0.0:
        raise ValueError("closed-form Taylor maps need a zero spiral angle")
    r = lat.r_axis()
    th = lat.fi_axis()
    d = lat.flutter_derivs(th, 2 * max_order + 1)
    gc = _g_coeffs(lat.k, max_order)
    shape = (max_order + 1, r.size, th.size)
    bz = np.empty(shape, dtype)
    br = np.empty(shape, dtype)
    bfi = np.empty(shape, dtype)
    for n in range(max_order + 1):
        g = np.zeros(th.size)
        gp = np.zeros(th.size)
        for i, c in enumerate(gc[n]):
            g += c * d[2 * i]
            gp += c * d[2 * i + 1]
        c2n = (-1) ** n / math.factorial(2 * n)
        c2n1 = (-1) ** n / math.factorial(2 * n + 1)
        rk = lat.B0 * r ** (lat.k - 2 * n)
        rk1 = lat.B0 * r ** (lat.k - 2 * n - 1)
        bz[n] = c2n * np.outer(rk, g)
        br[n] = c2n1 * (lat.k - 2 * n) * np.outer(rk1, g)
        bfi[n] = c2n1 * np.outer(rk1, gp)
    return bz, br, bfi


# ---------------------------------------------------------------------------
# Closed orbits (SEO)
# ---------------------------------------------------------------------------
def _propagate(lat, brho, y0, n_theta, record=False):
    """RK4 in theta over one sector for y = (r, alpha), alpha being the angle
    of the velocity to the azimuthal direction.  The tracker uses
    B_cart = -B_cyl (src/FFAG_MathTools.py:2011-2014), which makes a positive
    Bz bend a counter-clockwise proton towards the centre:
    dr/dtheta = r tan(alpha), dalpha/dtheta = 1 - Bz r / (Brho cos(alpha)), alpha
    being measured so that v_r = v sin(alpha) and v_phi = v cos(alpha)."""
    period = 2 * np.pi / lat.nsector
    h = period / n_theta
    # the flutter is sampled on the half-step grid once (it depends on theta only)
    fl = lat.flutter(np.arange(2 * n_theta + 1) * (h / 2))
    b0, kk = lat.B0, lat.k
    r, al = float(y0[0]), float(y0[1])
    rec = [(r, al)] if record else None

    def rhs(r_, a_, f_):
        return r_ * math.tan(a_), 1.0 - b0 * r_ ** kk * f_ / brho * r_ / math.cos(a_)

    for i in range(n_theta):
        f0, fh, f1 = fl[2 * i], fl[2 * i + 1], fl[2 * i + 2]
        k1r, k1a = rhs(r, al, f0)
        k2r, k2a = rhs(r + h / 2 * k1r, al + h / 2 * k1a, fh)
        k3r, k3a = rhs(r + h / 2 * k2r, al + h / 2 * k2a, fh)
        k4r, k4a = rhs(r + h * k3r, al + h * k3a, f1)
        r += h / 6 * (k1r + 2 * k2r + 2 * k3r + k4r)
        al += h / 6 * (k1a + 2 * k2a + 2 * k3a + k4a)
        if record:
            rec.append((r, al))
    y = np.array([r, al])
    return (y, np.array(rec)) if record else y


def closed_orbit(lat: ScalingFFAG, ek_mev: float, n_theta: int = 2000):
    """Closed orbit over one sector by Newton iteration on (r, alpha)."""
    brho = ek_to_p(ek_mev) / Q_E
    if lat.B0 is None:
        raise ValueError("B0 not set")
    # scaling-law guess
    y = np.array([lat.r0 * (brho / (lat.B0 * lat.r0 ** (lat.k + 1) * lat.mean_flutter())) ** (1 / (lat.k + 1)), 0.0])
    for _ in range(30):
        f0 = _propagate(lat, brho, y, n_theta) - y
        if np.max(np.abs(f0)) < 1e-13:
            break
        jac = np.zeros((2, 2))
        for j, eps in enumerate((1e-6, 1e-7)):
            dy = np.zeros(2)
            dy[j] = eps
            jac[:, j] = ((_propagate(lat, brho, y + dy, n_theta) - (y + dy)) - f0) / eps
        y = y - np.linalg.solve(jac, f0)
    _, rec = _propagate(lat, brho, y, n_theta, record=True)
    theta = np.linspace(0, 2 * np.pi / lat.nsector, n_theta + 1)
    r, al = rec[:, 0], rec[:, 1]
    # path length over one sector, times nsector
    ds = r / np.cos(al)
    per_sector = np.trapezoid(ds, theta)
    perimeter = per_sector * lat.nsector
    v = ek_to_v(ek_mev)
    return {"theta": theta, "r": r, "alpha": al, "pr": np.tan(al),
            "perimeter": perimeter, "T0": perimeter / v}


@dataclass
class SEOTables:
    """Closed-orbit tables in the layout the tracker consumes
    (``FFAG_Bunch.SEO_*`` attributes set at examples/main_track.py:333-337)."""
    Ek_axis: np.ndarray
    fi_axis: np.ndarray
    r_matrix: np.ndarray
    pr_matrix: np.ndarray
    perimeter: np.ndarray
    T0: np.ndarray
    info: np.ndarray = field(default=None)


def seo_tables(lat: ScalingFFAG, ek_list, n_fi: int | None = None, n_theta: int = 2000) -> SEOTables:
    """r0(Ek, phi), pr0(Ek, phi) over phi in [0, 2pi] (uniform axis of
    ``int(2 pi / 0.001)`` points, src/FFAG_track.py:622-623), perimeter and
    revolution period per energy."""
    ek_list = np.asarray(ek_list, dtype=np.float64)
    if n_fi is None:
        n_fi = int(2 * np.pi / 0.001)
    fi_axis = np.linspace(0, 2 * np.pi, n_fi)
    period = 2 * np.pi / lat.nsector
    rm = np.zeros((ek_list.size, n_fi))
    pm = np.zeros((ek_list.size, n_fi))
    per = np.zeros(ek_list.size)
    t0 = np.zeros(ek_list.size)
    for i, ek in enumerate(ek_list):
        co = closed_orbit(lat, ek, n_theta)
        th_mod = np.mod(fi_axis, period)
        rm[i] = np.interp(th_mod, co["theta"], co["r"])
        pm[i] = np.interp(th_mod, co["theta"], co["pr"])
        per[i] = co["perimeter"]
        t0[i] = co["T0"]
    info = np.zeros((ek_list.size, 9))
    info[:, 0] = np.arange(ek_list.size)
    info[:, 1] = ek_list
    info[:, 4] = 1.0 / t0
    info[:, 6] = rm[:, 0]
    info[:, 7] = pm[:, 0]
    info[:, 8] = per
    return SEOTables(ek_list, fi_axis, rm, pm, per, t0, info)


def write_seo(folder, seo: SEOTables):
    """Write SEO_ini/SEO_r/SEO_pr in the column layout main_track.py reads."""
    os.makedirs(folder, exist_ok=True)
    np.savetxt(os.path.join(folder, "SEO_ini.txt"), seo.info,
               header="idx Ek - - f_rev - r0 pr0 perimeter\nunits: - MeV - - Hz - m - m")
    for name, mat in (("SEO_r.txt", seo.r_matrix), ("SEO_pr.txt", seo.pr_matrix)):
        tab = np.zeros((mat.shape[0] + 1, mat.shape[1] + 1))
        tab[0, 1:] = seo.fi_axis
        tab[1:, 0] = seo.Ek_axis
        tab[1:, 1:] = mat
        np.savetxt(os.path.join(folder, name), tab, header="first row: fi axis; first column: Ek")
    with open(os.path.join(folder, "config_SEO.json"), "w") as f:
        json.dump({"BHarmonics": {"enable": False, "harmonics": []}}, f, indent=2)


def write_twiss(folder, seo: SEOTables, beta_r, beta_z, alpha_r=0.0, alpha_z=0.0):
    """BetaFuncR/Z.txt and AlphaFuncR/Z.txt (first column Ek, first row the
    azimuth axis) holding the given values at every (Ek, phi): the Twiss
    inputs the bunch generator reads at the injection point."""
    os.makedirs(folder, exist_ok=True)
    for name, v in (("BetaFuncR.txt", beta_r), ("BetaFuncZ.txt", beta_z), ("AlphaFuncR.txt", alpha_r),
                    ("AlphaFuncZ.txt", alpha_z)):
        tab = np.zeros((seo.Ek_axis.size + 1, seo.fi_axis.size + 1))
        tab[0, 1:] = seo.fi_axis
        tab[1:, 0] = seo.Ek_axis
        tab[1:, 1:] = v
        np.savetxt(os.path.join(folder, name), tab, header="first row: fi axis; first column: Ek")


def write_bmap(folder, lat: ScalingFFAG, max_order: int):
    """config_Bmap.json + per-order npz stacks, as FFAG_Bfield_analytical loads."""
    os.makedirs(folder, exist_ok=True)
    lat.folder = folder
    with open(os.path.join(folder, "config_Bmap.json"), "w") as f:
        json.dump(lat.to_config(), f, indent=2)
    bz, br, bfi = taylor_maps(lat, max_order)
    for n in range(max_order + 1):
        np.savez(os.path.join(folder, f"Bz_order_{n}.npz"), Bz_Taylor_coeff=bz[n])
        np.savez(os.path.join(folder, f"Br_order_{n}.npz"), Br_Taylor_coeff=br[n])
        np.savez(os.path.join(folder, f"Bfi_order_{n}.npz"), Bfi_Taylor_coeff=bfi[n])


# ---------------------------------------------------------------------------
# RF frequency curve (Emap)
# ---------------------------------------------------------------------------
def write_emap(path, *, gap_azimuth_deg=(0.0, 180.0), harmonic=1, v0_volt=100e3,
               f_start=None, f_end=None, t_start=-200e-9, t_end=1e-3, n_t=4001,
               gap_width=0.1, e_rmin=8.0, e_rmax=9.3, acc_phi_deg=0.0,
               ek_mev=300.0):
    """Write a fixed-gap RF file: header keys of examples/main_Emap.py:754-787,
    4 columns time [s], f_RF [Hz], V0 [V], Ek [MeV] on a uniform time grid, the
    frequency ramped linearly from ``f_start`` to ``f_end``."""
    ngap = len(gap_azimuth_deg)
    rf_shift = [-360.0 * harmonic * i / ngap for i in range(ngap)]
    t = np.linspace(t_start, t_end, n_t)
    f = np.linspace(f_start, f_end, n_t)
    v = np.full(n_t, v0_volt)
    e = np.full(n_t, ek_mev)
    hdr = [
        "gap_model = fixed",
        f"Ngap = {ngap}",
        f"acc_voltage_paint = {v0_volt}",
        f"acc_voltage_acc = {v0_volt}",
        f"acc_phi_paint = {float(acc_phi_deg)}",
        f"acc_phi_acc = {float(acc_phi_deg)}",
        f"gap_azimuth = {[float(g) for g in gap_azimuth_deg]}",
        f"rf_shift = {[float(s) for s in rf_shift]}",
        f"gap_width = {float(gap_width)}",
        f"E_rmin = {float(e_rmin)}",
        f"E_rmax = {float(e_rmax)}",
        f"harmonic = {int(harmonic)}",
        f"dt_uniform = {(t[1] - t[0]):.6e}",
        "#Time (s)        Frequency (Hz)    V0 (V)    Energy (MeV)",
    ]
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "w") as fh:
        fh.write("\n".join(hdr) + "\n")
        np.savetxt(fh, np.column_stack([t, f, v, e]), fmt="%.12e", delimiter="\t")


def write_spiral_emap(path, *, gap_azimuth_deg=(0.0, 180.0), harmonic=1, v0_volt=100e3,
                      f_start=None, f_end=None, t_start=-200e-9, t_end=1e-3, n_t=4001,
                      gap_width=0.05, spiral_deg=35.0, r_ref=8.5, rmin=8.0, rmax=9.3, nr=131, nphi=241,
                      acc_phi_deg=0.0, ek_mev=300.0):
    """Write a spiral-gap RF file (``gap_model = spiral``, header keys of
    examples/main_Emap.py:754-787) and its three maps next to it.

    The gap follows the spiral line phi_c(r) = tan(spiral) ln(r / r_ref) in
    the local azimuth; across it the accelerating field is a Gaussian of rms
    width gap_width / 2 whose integral over the normal distance is
    gap_width (so the crossing gains V0), pointing along the line's normal
    (Er, Ephi) = (-sin, cos)(spiral).  Ez / z = -div_perp E keeps the field
    divergence-free to first order in z.  Maps are per unit V0 / gap_width."""
    ngap = len(gap_azimuth_deg)
    rf_shift = [-360.0 * harmonic * i / ngap for i in range(ngap)]
    ta = np.tan(np.deg2rad(spiral_deg))
    r_axis = np.linspace(rmin, rmax, nr)
    pc = ta * np.log(r_axis / r_ref)
    margin = 4.0 * gap_width / rmin
    phi_axis = np.linspace(pc.min() - margin, pc.max() + margin, nphi)
    R, PHI = np.meshgrid(r_axis, phi_axis, indexing="ij")
    ca, sa = np.cos(np.deg2rad(spiral_deg)), np.sin(np.deg2rad(spiral_deg))
    d = (PHI - ta * np.log(R / r_ref)) * R * ca               # normal distance from the line
    sig = 0.5 * gap_width
    prof = np.exp(-0.5 * (d / sig) ** 2) * gap_width / (np.sqrt(2.0 * np.pi) * sig)
    er, ephi = -sa * prof, ca * prof
    dr, dphi = r_axis[1] - r_axis[0], phi_axis[1] - phi_axis[0]
    ezc = -(np.gradient(R * er, dr, axis=0) / R + np.gradient(ephi, dphi, axis=1) / R)
    d0 = os.path.dirname(path)
    if d0:
        os.makedirs(d0, exist_ok=True)
    names = {"Er_map_file": "Er_coef.txt", "Ez_map_file": "Ez_coef.txt", "Ephi_map_file": "Ephi_coef.txt"}
    for key, arr in (("Er_map_file", er), ("Ez_map_file", ezc), ("Ephi_map_file", ephi)):
        tab = np.zeros((nr + 1, nphi + 1))
        tab[0, 1:] = phi_axis
        tab[1:, 0] = r_axis
        tab[1:, 1:] = arr
        np.savetxt(os.path.join(d0, names[key]), tab, fmt="%.17e")
    t = np.linspace(t_start, t_end, n_t)
    f = np.linspace(f_start, f_end, n_t)
    ek_nodes = [ek_mev - 10.0, ek_mev + 10.0]
    hdr = [
        "gap_model = spiral",
        f"Ngap = {ngap}",
        f"acc_voltage_paint = {v0_volt}",
        f"acc_voltage_acc = {v0_volt}",
        f"acc_phi_paint = {float(acc_phi_deg)}",
        f"acc_phi_acc = {float(acc_phi_deg)}",
        f"gap_azimuth = {[float(g) for g in gap_azimuth_deg]}",
        f"rf_shift = {[float(s) for s in rf_shift]}",
        f"gap_width = {float(gap_width)}",
        f"E_rmin = {float(rmin)}",
        f"E_rmax = {float(rmax)}",
        f"harmonic = {int(harmonic)}",
        f"spiral_alpha_deg = {float(spiral_deg)}",
        f"spiral_r_ref = {float(r_ref)}",
        "spiral_s_ref_deg = 0.0",
        "spiral_gap_node_count = 2",
        f"Ek_gap = {ek_nodes}",
        f"r_gap = {[float(r_ref), float(r_ref)]}",
        "phi_gap_deg = [0.0, 0.0]",
        "phi_gap_unwrapped_deg = [0.0, 0.0]",
        f"dt_uniform = {(t[1] - t[0]):.6e}",
        f"Er_map_file = {names['Er_map_file']}",
        f"Ez_map_file = {names['Ez_map_file']}",
        f"Ephi_map_file = {names['Ephi_map_file']}",
        "#Time (s)        Frequency (Hz)    V0 (V)    Energy (MeV)",
    ]
    with open(path, "w") as fh:
        fh.write("\n".join(hdr) + "\n")
        np.savetxt(fh, np.column_stack([t, f, np.full(n_t, v0_volt), np.full(n_t, ek_mev)]), fmt="%.12e",
                   delimiter="\t")


# ---------------------------------------------------------------------------
# Particle distribution
# ---------------------------------------------------------------------------
def gaussian_bunch(lat: ScalingFFAG, seo: SEOTables, n: int, seed: int = 0, *,
                   ek_mev: float = 300.0, emit_rms=10e-6, beta_r=None, beta_z=None,
                   dpp_sigma=1e-3, inj_t_ns=0.0, fi_spread=2 * np.pi):
    """Seeded Gaussian (x, x', y, y') bunch with rms unnormalised emittance
    ``emit_rms`` about the closed orbit, Gaussian dp/p and a uniform azimuth
    spread over ``fi_spread``; every particle shares injection time
    ``inj_t_ns``.  Returns the reference's 16-column (r, pr, z, pz, fi, Ek, ...)
    matrix (src/FFAG_Bunch.py:23-62).  Betas default to the smooth-focusing
    estimates R/Q with Q_r = sqrt(1+k) and Q_z = 1.5."""
    rng = np.random.default_rng(seed)
    if beta_r is None:
        beta_r = lat.r0 / math.sqrt(1.0 + lat.k)
    if beta_z is None:
        beta_z = lat.r0 / 1.5
    x = rng.normal(0.0, math.sqrt(emit_rms * beta_r), n)
    xp = rng.normal(0.0, math.sqrt(emit_rms / beta_r), n)
    y = rng.normal(0.0, math.sqrt(emit_rms * beta_z), n)
    yp = rng.normal(0.0, math.sqrt(emit_rms / beta_z), n)
    dpp = rng.normal(0.0, dpp_sigma, n)
    fi = rng.uniform(0.0, fi_spread, n)
    p0 = ek_to_p(ek_mev)
    p = p0 * (1.0 + dpp)
    ek = (np.sqrt((p * C_LIGHT) ** 2 + E0_J ** 2) - E0_J) / Q_E / 1e6
    # closed orbit at the reference energy, scaled to each momentum
    j = int(np.argmin(np.abs(seo.Ek_axis - ek_mev)))
    fim = np.mod(fi, 2 * np.pi)
    rco = np.interp(fim, seo.fi_axis, seo.r_matrix[j]) * (1.0 + dpp) ** (1.0 / (lat.k + 1.0))
    prco = np.interp(fim, seo.fi_axis, seo.pr_matrix[j])
    out = np.zeros((n, 16))
    out[:, 0] = rco + x
    out[:, 1] = np.arctan(prco) + xp
    out[:, 2] = y
    out[:, 3] = yp
    out[:, 4] = fi
    out[:, 5] = ek
    out[:, 6] = inj_t_ns
    out[:, 13] = 0
    out[:, 14] = np.arange(n)
    out[:, 15] = np.arange(n)
    return out
