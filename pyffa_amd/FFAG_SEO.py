"""Closed orbits (SEO), tunes and Twiss functions on the GPU.

Same entry points and outputs as the reference:

* ``FFAG_SearchSEO(GlobalParas)`` (src/FFAG_track.py:1142-1586):
  ``TrackAParticle``, ``SearchSEOUsingInitialEkVect``, ``GetQrQz``,
  ``GetFreq`` and ``SearchSEOsControllerVect(EkRange)``, which writes
  ``SEO_ini.txt`` (17 columns: index, Ek, Qr, Qz, f_rev, mean r, r0, pr0,
  perimeter, r and pr after a turn and at 90/180/270 deg), ``SEO_r.txt``
  and ``SEO_pr.txt`` into ``GlobalParas.Bmap.BmapFoldName /
  GlobalParas.SEO_SaveFold``;
* ``FFAG_BetaFuncCalc().CalcBetaFunc(SEOPATHName, Ek_arr)``
  (src/FFAG_BetaFunc.py:45-216), which writes ``BetaFuncR/Z.txt`` and
  ``AlphaFuncR/Z.txt`` (first row the azimuth in degrees over one sector).

All tracking runs in ``ffa_orbit_track`` (``k_orbit_rk4``): RK4 in azimuth
of the reference's closed-orbit equations, one particle per thread, in the
order-0 field plus harmonics.  The searches are batched over energies: every
iteration of the midpoint / Newton search, every tune matrix and the Twiss
two-cell track is one launch for all energies at once (the reference loops
over energies and tracks one particle at a time).
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import types

import numpy as np

from . import _lib
from .FFAG_Field import build_machine
from .FFAG_ParasAndConversion import FFAG_ConversionTools

PR0_TRIALS = [0.00, 0.02, -0.02, 0.04, -0.04, 0.06, -0.06, 0.08, -0.08, 0.10, -0.10, 0.12, -0.12, 0.14, -0.14]
SEO_TITLES = ["Ek_index", "Ek_value(MeV)", "Qr", "Qz", "OrbitFreq(Hz)", "MeanR(m)", "r0(m)", "pr0", "Perimeter(m)",
              "r360(m)", "pr360", "r90(m)", "pr90", "r180(m)", "pr180", "r270(m)", "pr270"]


def _grid(t0, t1, h):
    """The rk4_solve grid: linspace(t0, t1, int((t1 - t0) / h)); returns
    (points, linspace step, RK4 step = points[1] - points[0])."""
    steps = int((t1 - t0) / h)
    tp = np.linspace(t0, t1, steps)
    return tp, (t1 - t0) / (steps - 1), tp[1] - tp[0]


class OrbitTracker:
    """GPU tracking of (r, pr, z, pz) in azimuth for the field of
    ``GlobalParas`` (order 0, plus the harmonic errors unless
    ``harmonics`` is False)."""

    def __init__(self, GlobalParas, device=None, harmonics=True):
        import torch
        self.torch = torch
        if not torch.cuda.is_available():
            raise _lib.FfaError("the closed-orbit search runs on the GPU")
        self.gp = GlobalParas
        self.dev = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        self.ctx = _lib.Context(self.dev.index, torch.cuda.current_stream(self.dev).cuda_stream)
        light = types.SimpleNamespace(Bmap=GlobalParas.Bmap,
                                      BHarmonics=GlobalParas.BHarmonics if harmonics else None, Emap=None,
                                      Aperture_enable=False, q=GlobalParas.q, c=GlobalParas.c, m0=GlobalParas.m0,
                                      E0=GlobalParas.E0)
        m = build_machine(light, None, False)
        m.max_order = 0
        m.taylor = None
        self.ctx.check(self.ctx.lib.ffa_set_machine(self.ctx.h, ctypes.byref(m)), "ffa_set_machine")
        self.q = GlobalParas.q
        self.conv = FFAG_ConversionTools()

    def q_over_p(self, ek):
        return np.array([self.q / self.conv.Ek2P(float(e)) for e in np.atleast_1d(ek)], dtype=np.float64)

    def track(self, X, ek, t_start, t_end, h, record=False):
        """States (n, 4) at t_end, or (grid, trajectory (points, n, 4)) with
        ``record``; ek (n,) in MeV."""
        t = self.torch
        X = np.ascontiguousarray(np.asarray(X, dtype=np.float64).reshape(-1, 4))
        n = X.shape[0]
        tp, step, hh = _grid(t_start, t_end, h)
        state = t.from_numpy(np.ascontiguousarray(X.T)).to(self.dev)
        qp = t.from_numpy(self.q_over_p(ek)).to(self.dev)
        traj = t.empty((len(tp), 4, n), dtype=t.float64, device=self.dev) if record else None
        o = _lib.FfaOrbit(state=state.data_ptr(), q_over_p=qp.data_ptr(), n=n, t0=float(t_start), t_step=step, h=hh,
                          nsteps=len(tp) - 1, pad=0, traj=traj.data_ptr() if record else None)
        self.ctx.check(self.ctx.lib.ffa_orbit_track(self.ctx.h, ctypes.byref(o)), "ffa_orbit_track")
        if record:
            return tp, traj.cpu().numpy().transpose(0, 2, 1)
        return state.cpu().numpy().T.copy()


class FFAG_SearchSEO:
    def __init__(self, GlobalParas, device=None):
        self.GlobalParas = GlobalParas
        self.tracker = OrbitTracker(GlobalParas, device)

    # -- single-particle views of the reference API ---------------------------
    def TrackAParticle(self, r0, pr0, Ek_value, t_end, h):
        """(t_points, r_points) with rows (r, pr, z, pz, P, index)."""
        tp, tr = self.tracker.track([[r0, pr0, 0.0, 0.0]], [Ek_value], 0.0, t_end, h, record=True)
        rp = np.zeros((len(tp), 6))
        rp[:, :4] = tr[:, 0, :]
        rp[:, 4] = FFAG_ConversionTools().Ek2P(Ek_value)
        return tp, rp

    def mean_field_radius(self, ek):
        """Radius where P / q = r <Bz>(r), <Bz> the azimuthal mean of the
        order-0 field over the sector (the reference's bisection on its
        mean-field curve f1)."""
        bm = self.GlobalParas.Bmap
        th = np.linspace(0.0, 2 * np.pi / bm.Nsectors, 721)
        brho = np.array([FFAG_ConversionTools().Ek2P(float(e)) for e in np.atleast_1d(ek)]) / self.GlobalParas.q
        from .FFAG_Field import mean_order0_bz
        lo, hi = np.full(brho.size, bm.r_min), np.full(brho.size, bm.r_max)

        def f(r):
            return r - brho / mean_order0_bz(bm, r, th)
        flo = f(lo)
        for _ in range(100):
            mid = (lo + hi) / 2
            fm = f(mid)
            left = fm * flo < 0
            hi = np.where(left, mid, hi)
            lo = np.where(left, lo, mid)
            flo = np.where(left, flo, fm)
        return (lo + hi) / 2

    def search(self, ek_list, h=0.001, delta_r0_ini=0.0, full_turn_newton=None):
        """(r0, pr0, r0_analytical) arrays: the midpoint iteration over one
        sector from the mean-field radius with the pr0 trials on NaN, then the
        finite-difference Newton iteration (over the full turn when harmonics
        are on), all energies per launch."""
        ek = np.atleast_1d(np.asarray(ek_list, dtype=np.float64))
        n = ek.size
        bm = self.GlobalParas.Bmap
        sector = np.pi * 2 / bm.Nsectors
        r_an = self.mean_field_radius(ek)
        r0 = r_an + delta_r0_ini
        trial = np.zeros(n, dtype=np.int64)
        pr0 = np.full(n, PR0_TRIALS[0])
        done = np.zeros(n, bool)
        it = np.zeros(n, dtype=np.int64)
        while not done.all():
            act = np.flatnonzero(~done)
            X = np.zeros((act.size, 4))
            X[:, 0], X[:, 1] = r0[act], pr0[act]
            E = self.tracker.track(X, ek[act], 0.0, sector, h)
            bad = np.isnan(E[:, 0]) | np.isnan(E[:, 1])
            conv = ~bad & (np.abs(X[:, 0] - E[:, 0]) < 1e-5) & (np.abs(X[:, 1] - E[:, 1]) < 1e-5)
            nr, npr = (X[:, 0] + E[:, 0]) / 2, (X[:, 1] + E[:, 1]) / 2
            for j, i in enumerate(act):
                it[i] += 1
                if conv[j]:
                    # the reference updates to the midpoint before it tests convergence
                    r0[i], pr0[i] = nr[j], npr[j]
                    done[i] = True
                elif bad[j] or it[i] >= 200:
                    trial[i] += 1
                    if trial[i] >= len(PR0_TRIALS):
                        raise RuntimeError(f"No close orbit solution found for {ek[i]}MeV.")
                    r0[i], pr0[i], it[i] = r_an[i] + delta_r0_ini, PR0_TRIALS[trial[i]], 0
                else:
                    r0[i], pr0[i] = nr[j], npr[j]
        harm = self.GlobalParas.BHarmonics is not None and self.GlobalParas.BHarmonics.enable_flag
        t_end = np.pi * 2 if (harm if full_turn_newton is None else full_turn_newton) else sector
        active = np.ones(n, bool)
        for _ in range(25):
            act = np.flatnonzero(active)
            if act.size == 0:
                break
            dr, dp = r0[act] * 1e-4, 1e-3
            X = np.zeros((3 * act.size, 4))
            X[0::3, 0], X[0::3, 1] = r0[act], pr0[act]
            X[1::3, 0], X[1::3, 1] = r0[act] + dr, pr0[act]
            X[2::3, 0], X[2::3, 1] = r0[act], pr0[act] + dp
            E = self.tracker.track(X, np.repeat(ek[act], 3), 0.0, t_end, h)
            r0f, p0f = E[0::3, 0], E[0::3, 1]
            a11, a12 = (E[1::3, 0] - r0f) / dr, (E[2::3, 0] - r0f) / dp
            a21, a22 = (E[1::3, 1] - p0f) / dr, (E[2::3, 1] - p0f) / dp
            a11p, a22p = a11 - 1, a22 - 1
            det = a11p * a22p - a12 * a21
            stop = np.abs((r0f - r0[act]) / r0[act]) < 1e-8
            re = r0[act] + a22p / det * (r0[act] - r0f) - a12 / det * (pr0[act] - p0f)
            pe = pr0[act] + a11p / det * (pr0[act] - p0f) - a21 / det * (r0[act] - r0f)
            r0[act] = np.where(stop, r0[act], re)
            pr0[act] = np.where(stop, pr0[act], pe)
            active[act[stop]] = False
        return r0, pr0, r_an

    def SearchSEOUsingInitialEkVect(self, Ek_value, h=0.001, delta_r0_ini=0.0, verbose=False):
        r0, pr0, r_an = self.search([Ek_value], h, delta_r0_ini)
        return float(r0[0]), float(pr0[0]), float(r_an[0])

    def GetFreq(self, fi_SEO, r_SEO, pr_SEO, Ek_value):
        rp = pr_SEO / (np.sqrt(1 - pr_SEO ** 2)) * r_SEO
        dl = np.sqrt(r_SEO ** 2 + rp ** 2) * (fi_SEO[2] - fi_SEO[1])
        perimeter = sum(dl[:-1]) * self.GlobalParas.Bmap.Nsectors
        return 1 / (perimeter / FFAG_ConversionTools().Ek2v(Ek_value)), perimeter

    def tunes(self, r0, pr0, ek, h=0.001):
        """Arrays (Qr, Qz, f_rev, mean r, perimeter) from one-sector
        matrices of 5 particles per energy (GetQrQz)."""
        r0, pr0, ek = (np.atleast_1d(np.asarray(a, dtype=np.float64)) for a in (r0, pr0, ek))
        n = ek.size
        N = self.GlobalParas.Bmap.Nsectors
        drr, dpp, dzz, dpz = 1e-6, 1e-6, 1e-7, 1e-7
        X = np.zeros((5 * n, 4))
        X[:, 0], X[:, 1] = np.repeat(r0, 5), np.repeat(pr0, 5)
        X[1::5, 0] += drr
        X[2::5, 1] += dpp
        X[3::5, 2] += dzz
        X[4::5, 3] += dpz
        tp, tr = self.tracker.track(X, np.repeat(ek, 5), 0.0, np.pi * 2 / N, h, record=True)
        E = tr[-1]
        a11 = (E[1::5, 0] - E[0::5, 0]) / drr
        a22 = (E[2::5, 1] - E[0::5, 1]) / dpp
        b11 = (E[3::5, 2] - 0.0) / dzz
        b22 = (E[4::5, 3] - 0.0) / dpz
        qr = np.arccos((a11 + a22) / 2) / (2 * np.pi) * N
        qz = np.arccos((b11 + b22) / 2) / (2 * np.pi) * N
        f, per, mr = np.zeros(n), np.zeros(n), np.zeros(n)
        for i in range(n):
            f[i], per[i] = self.GetFreq(tp, tr[:, 5 * i, 0], tr[:, 5 * i, 1], ek[i])
            mr[i] = np.mean(tr[:, 5 * i, 0])
        return qr, qz, f, mr, per

    def GetQrQz(self, r0, pr0, Ek_value, h):
        qr, qz, f, mr, per = self.tunes([r0], [pr0], [Ek_value], h)
        return float(qr[0]), float(qz[0]), float(f[0]), float(mr[0]), float(per[0])

    def SearchSEOsControllerVect(self, EkRange, SavedataFlag=True):
        """Closed orbits, tunes and the r(phi), pr(phi) tables for every
        energy; writes the three SEO files and returns their folder."""
        ek = np.asarray(EkRange, dtype=np.float64)
        h = 0.001
        r0, pr0, _ = self.search(ek, h)
        X = np.zeros((ek.size, 4))
        X[:, 0], X[:, 1] = r0, pr0
        tp, tr = self.tracker.track(X, ek, 0.0, np.pi * 2, 0.001, record=True)
        qr, qz, f, mr, per = self.tunes(r0, pr0, ek, h)
        data = np.zeros((ek.size, 17))
        for i in range(ek.size):
            rr, pp = tr[:, i, 0], tr[:, i, 1]
            at = [float(np.interp(np.deg2rad(a), tp, rr)) for a in (90.0, 180.0, 270.0)]
            ap = [float(np.interp(np.deg2rad(a), tp, pp)) for a in (90.0, 180.0, 270.0)]
            data[i] = (i, ek[i], qr[i], qz[i], f[i], mr[i], r0[i], pr0[i], per[i], rr[-1], pp[-1],
                       at[0], ap[0], at[1], ap[1], at[2], ap[2])
        self.last = dict(data=data, fi=tp, r=tr[:, :, 0].T.copy(), pr=tr[:, :, 1].T.copy())
        if not SavedataFlag:
            return None
        bm = self.GlobalParas.Bmap
        folder = os.path.join(bm.BmapFoldName, getattr(self.GlobalParas, "SEO_SaveFold", "resultsSEO"))
        os.makedirs(folder, exist_ok=True)
        write_seo_files(folder, data, tp, self.last["r"], self.last["pr"], bm.BmapFoldName, h)
        cfg = os.path.join(folder, "config_SEO.json")
        if not os.path.exists(cfg):
            bh = self.GlobalParas.BHarmonics
            with open(cfg, "w") as fh:
                json.dump({"BHarmonics": {"enable": bool(bh.enable_flag) if bh is not None else False,
                                          "harmonics": list(bh.harmonics_cfg) if bh is not None else []}}, fh,
                          indent=2)
        return folder


def write_seo_files(folder, data, fi, r_tab, pr_tab, bmap_folder, h):
    """SEO_ini.txt (a title line over left-aligned columns; Qr..MeanR rounded
    to 3 decimals, r0 onwards to 8), SEO_r.txt and SEO_pr.txt (a unit line,
    then the azimuth row and one row per energy led by Ek)."""
    d = np.array(data, dtype=float)
    d[:, 2:6] = np.round(d[:, 2:6], 3)
    d[:, 6:] = np.round(d[:, 6:], 8)
    rows = [SEO_TITLES] + d.tolist()
    widths = [max(len(str(row[i])) for row in rows) for i in range(len(SEO_TITLES))]
    lines = ["Bmap foldname=" + str(bmap_folder) + ", step h = %.6f rad" % (h,)]
    lines += ["  ".join(str(v).ljust(w) for v, w in zip(row, widths)) for row in rows]
    with open(os.path.join(folder, "SEO_ini.txt"), "w") as f:
        f.write("\n".join(lines))
    ek = d[:, 1]
    for name, tab, unit in (("SEO_r.txt", r_tab, "unitR=m"), ("SEO_pr.txt", pr_tab, "unitPr=1")):
        m = np.zeros((ek.size + 1, len(fi) + 1))
        m[0, 0], m[0, 1:] = ek[0], fi
        m[1:, 0], m[1:, 1:] = ek, tab
        with open(os.path.join(folder, name), "w") as f:
            f.write(f"unitEk=MeV, unitFi=deg, {unit}\n")
            np.savetxt(f, m)


class FFAG_BetaFuncCalc:
    """Twiss functions from one-sector transfer matrices
    (src/FFAG_BetaFunc.py:26-216)."""

    @staticmethod
    def find_coef(r1_s, pr1_s, r1_e, pr1_e, r2_s, pr2_s, r2_e, pr2_e):
        """(a11, a12, a21, a22) mapping two start vectors to two end vectors
        (Cramer's rule on the 4 x 4 system)."""
        A = np.array([[r1_s, pr1_s, 0, 0], [0, 0, r1_s, pr1_s], [r2_s, pr2_s, 0, 0], [0, 0, r2_s, pr2_s]])
        den = np.linalg.det(A)
        out = []
        for col in range(4):
            B = A.copy()
            B[:, col] = [r1_e, pr1_e, r2_e, pr2_e]
            out.append(np.linalg.det(B) / den)
        return tuple(out)

    def twiss(self, GlobalParas, r0, pr0, ek, beta_n=1000, h=0.001):
        """(fi0 [rad], beta_r, beta_z, alpha_r, alpha_z) with arrays
        (n_energies, beta_n) over one sector."""
        # the reference tracks the Twiss cells in a parameter set holding the
        # field map alone (src/FFAG_BetaFunc.py:76-80): no harmonic errors
        tracker = OrbitTracker(GlobalParas, harmonics=False)
        r0, pr0, ek = (np.atleast_1d(np.asarray(a, dtype=np.float64)) for a in (r0, pr0, ek))
        n = ek.size
        sector = np.pi * 2 / GlobalParas.Bmap.Nsectors
        d = 1e-7
        X = np.zeros((5 * n, 4))
        X[:, 0], X[:, 1] = np.repeat(r0, 5), np.repeat(pr0, 5)
        X[1::5, 0] += d
        X[2::5, 1] += d
        X[3::5, 2] += d
        X[4::5, 3] += d
        tp, tr = tracker.track(X, np.repeat(ek, 5), 0.0, 2 * sector, h, record=True)
        self._last_track = (tracker, tp, tr)
        fi0 = np.linspace(0, sector, beta_n)

        def at(t):
            # linear interpolation of every particle's state at azimuth t
            j = np.clip(np.searchsorted(tp, t, side="right") - 1, 0, len(tp) - 2)
            w = (t - tp[j]) / (tp[j + 1] - tp[j])
            return tr[j] * (1 - w) + tr[j + 1] * w
        out = np.zeros((4, n, beta_n))
        for k, f in enumerate(fi0):
            s, e = at(f), at(f + sector)
            for i in range(n):
                ds, de = s[5 * i + 1:5 * i + 5] - s[5 * i], e[5 * i + 1:5 * i + 5] - e[5 * i]
                a = self.find_coef(ds[0, 0], ds[0, 1], de[0, 0], de[0, 1], ds[1, 0], ds[1, 1], de[1, 0], de[1, 1])
                b = self.find_coef(ds[2, 2], ds[2, 3], de[2, 2], de[2, 3], ds[3, 2], ds[3, 3], de[3, 2], de[3, 3])
                for p, (m11, m12, m21, m22) in enumerate((a, b)):
                    c = (m11 + m22) / 2
                    sn = math.sqrt(1 - c ** 2)
                    if m12 * sn < 0:
                        sn = -sn
                    out[p, i, k] = m12 / sn
                    out[2 + p, i, k] = (m11 - m22) / sn / 2
        return fi0, out[0], out[1], out[2], out[3]

    def trajectory_fields(self):
        """Fields along the last ``twiss`` track, as the reference records
        them (src/FFAG_BetaFunc.py:82-84, 178-180): Bz of each energy's
        closed-orbit particle and Br, Bfi of its companion displaced in z,
        at every recorded azimuth.  Returns (azimuth [rad], Bz, Br, Bfi) with
        arrays (n_energies, points)."""
        if getattr(self, "_last_track", None) is None:
            raise RuntimeError("trajectory_fields() needs a twiss() track first")
        tracker, tp, tr = self._last_track
        t = tracker.torch
        n = tr.shape[1] // 5
        pts = np.concatenate([tr[:, 0::5, :], tr[:, 3::5, :]], axis=1)          # (points, 2n, 4)
        m = pts.shape[0] * pts.shape[1]
        dev = tracker.dev
        r = t.from_numpy(np.ascontiguousarray(pts[:, :, 0].T.ravel())).to(dev)
        z = t.from_numpy(np.ascontiguousarray(pts[:, :, 2].T.ravel())).to(dev)
        fi = t.from_numpy(np.ascontiguousarray(np.tile(np.asarray(tp, dtype=np.float64), 2 * n))).to(dev)
        out = t.empty(3 * m, dtype=t.float64, device=dev)
        lib = tracker.ctx.lib
        tracker.ctx.check(lib.ffa_bfield_eval(tracker.ctx.h, _lib._vp(r.data_ptr()), _lib._vp(fi.data_ptr()),
                                              _lib._vp(z.data_ptr()), m, _lib._vp(out.data_ptr()), 0, 0.0),
                          "ffa_bfield_eval")
        B = out.cpu().numpy().reshape(3, 2 * n, -1)                              # planes Bz, Bfi, Br
        return np.asarray(tp, dtype=np.float64), B[0, :n], B[2, n:], B[1, n:]

    def CalcBetaFunc(self, SEOPATHName, Ek_arr, GlobalParas):
        """Writes BetaFuncR/Z.txt and AlphaFuncR/Z.txt for the energies
        Ek_arr at the closed orbits of SEO_ini.txt (linear in Ek)."""
        from .FFAG_Distribution import LoadSEOParams
        ek = np.atleast_1d(np.asarray(Ek_arr, dtype=np.float64))
        r0, pr0 = np.zeros(ek.size), np.zeros(ek.size)
        for i, e in enumerate(ek):
            v, _, _ = LoadSEOParams(os.path.join(SEOPATHName, "SEO_ini.txt"), e)
            r0[i], pr0[i] = v[1], v[2]
        fi0, br, bz, ar, az = self.twiss(GlobalParas, r0, pr0, ek)
        head = np.insert(ek, 0, 0.0)
        for name, tab in (("BetaFuncR.txt", br), ("BetaFuncZ.txt", bz), ("AlphaFuncR.txt", ar),
                          ("AlphaFuncZ.txt", az)):
            m = np.column_stack((head, np.vstack((np.rad2deg(fi0), tab))))
            with open(os.path.join(SEOPATHName, name), "w") as f:
                f.write("unitEk=MeV, unitFi=deg\n")
                np.savetxt(f, m)
        # the trajectory fields; like the reference, the first row holds the
        # tracking azimuth as tracked (radians) under a "unitFi=deg" header
        tp, Bz, Br, Bf = self.trajectory_fields()
        for name, tab in (("Bz.txt", Bz), ("Br.txt", Br), ("Bf.txt", Bf)):
            m = np.column_stack((head, np.vstack((tp, tab))))
            with open(os.path.join(SEOPATHName, name), "w") as f:
                f.write("unitEk=MeV, unitFi=deg, unitB=T\n")
                np.savetxt(f, m)
        return fi0, br, bz, ar, az
