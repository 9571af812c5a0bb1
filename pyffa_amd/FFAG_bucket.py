"""Longitudinal RF bucket of the injection generator (src/FFAG_bucket.py).

The "match" longitudinal distribution of ``GenerateInitDistributionDtInjNew``
draws (arrival time, kinetic energy) pairs inside an RF bucket whose
projected length is the requested bunch length.  The model is the
reference's one-gap-per-turn longitudinal map written as a Hamiltonian flow
in q = zd (time offset from the synchronous particle, s) and p = dEk (MeV):

    dq/dt = A(p) = -(Tc - T(E0 + p)) / Tc
    dp/dt = B(q) = U0 (sin(phi_s + w0 q) - sin(phi_s)) / Tc / 1e6

(src/FFAG_bucket.py:27-44), with T(Ek) the revolution period of the closed
orbits (SEO_ini.txt columns Ek and f_rev, linear interpolation held at the
table ends), Tc = T(E0) and w0 = 2 pi f_rev(E0).  H(q, p) = -int B dq +
int A dp is conserved along the flow; the bucket is the region H >= H_sep
around the stable point (0, 0), which is H's maximum.

Host-side set-up code (numpy): a few dozen test trajectories integrated once
per generator call, not part of the tracking hot path.  Plotting is not
provided; ``show`` is accepted and ignored.
"""
import os

import numpy as np


def _linear_held(X, Y, xi):
    """Linear interpolation on the bracketing pair of a sorted axis, the end
    pair outside it (find_two_points_non_uniform_vect + linear form,
    src/FFAG_MathTools.py:59-82, 162-173)."""
    xi = np.asarray(xi, dtype=np.float64)
    i = np.clip(np.searchsorted(X, xi, side="right") - 1, 0, len(X) - 2)
    x0, x1, y0, y1 = X[i], X[i + 1], Y[i], Y[i + 1]
    return y0 + (xi - x0) / (x1 - x0) * (y1 - y0)


class FFAG_bucket:
    """RF bucket about the synchronous particle of energy Ek0 [MeV] with the
    synchronous phase fi0 [deg] and the energy gain amplitude U0 [eV] per
    turn (src/FFAG_bucket.py:51-123).  On construction 20 trajectories
    started on a ring about the unstable fixed point and 20 on the line from
    -0.8 |zd_u| to 0 are integrated forwards and backwards (RK4, step_t
    seconds, step_N steps) and classified as bounded or not; their extent
    sets the sampling box."""

    def __init__(self, Bmapfoldname, Ek0=300.0, fi0=30.0, U0=100e3, step_t=20e-9, step_N=15000):
        data = np.loadtxt(os.path.join(Bmapfoldname, "resultsSEO", "SEO_ini.txt"), skiprows=2, ndmin=2)
        ek, freq = data[:, 1].astype(np.float64), data[:, 4].astype(np.float64)
        order = np.argsort(ek)
        self.Ek_grid = np.ascontiguousarray(ek[order])
        self.T_grid = np.ascontiguousarray(1.0 / freq[order])
        self.E0 = float(Ek0)
        self.fi0 = float(np.deg2rad(fi0))
        self.U0 = float(U0)
        self.step_t = float(step_t)
        self.max_stepsN = int(step_N)
        self.Tc = float(_linear_held(self.Ek_grid, self.T_grid, self.E0))
        self.w0 = float(_linear_held(self.Ek_grid, freq[order], self.E0) * 2.0 * np.pi)
        # unstable fixed point: sin(phi_s + w0 q) = sin(phi_s) at phi = pi - phi_s
        self.zd_u = (np.pi - 2.0 * self.fi0) / self.w0
        th = np.linspace(0.0, 2.0 * np.pi, 20, endpoint=False)
        self.X_sep = np.column_stack((self.zd_u + 5.0e-9 * np.cos(th), 0.5 * np.sin(th)))
        z_ns = np.linspace(-0.8 * abs(self.zd_u) * 1e9, 0.0, 20)
        self.X_center = np.column_stack((z_ns * 1e-9, np.zeros(20)))
        self.X_in = np.vstack([self.X_sep, self.X_center])
        self.tr_fwd = self.rk4_solve_dt(self.dynamic, 0.0, self.X_in)
        self.stable_fwd, _ = self.classify_stability(self.tr_fwd)
        self.tr_rev = self.rk4_solve_dt(self.dynamic_neg, 0.0, self.X_in)
        self.stable_rev, _ = self.classify_stability(self.tr_rev)
        self.x_min_plot, self.x_max_plot, self.y_min_plot, self.y_max_plot = self._compute_axes_limits()
        self.x_min_sample, self.x_max_sample, self.y_min_sample, self.y_max_sample = self._compute_sample_limits()
        self.x_ns = self.y_Ek = self.q_grid = self.p_grid = None
        self.Hq_rel = self.Hp_rel = self.H00 = self.H_map = self.H_sep = None
        self._H_arr_center = self._L_mono_center = None

    # -- equations of motion (src/FFAG_bucket.py:27-44, 125-135) --
    def dynamic(self, t, X_in):
        X = np.asarray(X_in, dtype=np.float64)
        out = np.empty_like(X)
        T = np.interp(self.E0 + X[:, 1], self.Ek_grid, self.T_grid)
        out[:, 0] = -(self.Tc - T) / self.Tc
        out[:, 1] = self.U0 * (np.sin(self.fi0 + self.w0 * X[:, 0]) - np.sin(self.fi0)) / self.Tc / 1e6
        return out

    def dynamic_neg(self, t, X_in):
        return -self.dynamic(t, X_in)

    def rk4_step(self, func, t, r, h):
        k1 = h * func(t, r)
        k2 = h * func(t + 0.5 * h, r + 0.5 * k1)
        k3 = h * func(t + 0.5 * h, r + 0.5 * k2)
        k4 = h * func(t + h, r + k3)
        return (k1 + 2 * k2 + 2 * k3 + k4) / 6

    def rk4_solve_dt(self, func, t_start, r_start):
        """(step_N, P, 3) history of (t, zd, dEk) before each step."""
        r = np.array(r_start, dtype=np.float64)
        tr = np.zeros((self.max_stepsN, r.shape[0], 3))
        t = float(t_start)
        for i in range(self.max_stepsN):
            tr[i, :, 0] = t
            tr[i, :, 1:] = r
            r = r + self.rk4_step(func, t, r, self.step_t)
            t += self.step_t
        return tr

    # -- stability and ranges (src/FFAG_bucket.py:164-233) --
    @staticmethod
    def _wrap_pi(x):
        return (x + np.pi) % (2 * np.pi) - np.pi

    def classify_stability(self, tr_arr, growth_tol=0.2, frac_split=0.5):
        """A trajectory is bounded unless its phase offset w0 zd reaches pi or
        the largest |phase| or |dEk| of its second part exceeds that of its
        first part by more than growth_tol."""
        steps, P, _ = tr_arr.shape
        split = max(4, int(steps * frac_split))
        stable = np.ones(P, dtype=bool)
        reasons = [""] * P
        for pid in range(P):
            zd, de = tr_arr[:, pid, 1], tr_arr[:, pid, 2]
            if np.any(np.abs(self.w0 * zd) >= np.pi):
                stable[pid] = False
                reasons[pid] = "crossed |phi-fi0|=pi"
                continue
            ph = self._wrap_pi(self.fi0 + self.w0 * zd - self.fi0)
            p1 = max(np.max(np.abs(ph[:split])), 1e-12)
            p2 = np.max(np.abs(ph[split:])) if split < steps else p1
            e1 = max(np.max(np.abs(de[:split])), 1e-12)
            e2 = np.max(np.abs(de[split:])) if split < steps else e1
            if p2 / p1 > 1.0 + growth_tol or e2 / e1 > 1.0 + growth_tol:
                stable[pid] = False
                reasons[pid] = f"amplitude growth (R_phi={p2 / p1:.2f}, R_E={e2 / e1:.2f})"
            else:
                reasons[pid] = "bounded"
        return stable, reasons

    def _bounded_points(self):
        zs, es = [], []
        for tr, mask in ((self.tr_fwd, self.stable_fwd), (self.tr_rev, self.stable_rev)):
            if mask.any():
                zs.append((tr[:, mask, 1] * 1e9).ravel())
                es.append((tr[:, mask, 2] + self.E0).ravel())
        return zs, es

    def _compute_sample_limits(self):
        zs, es = self._bounded_points()
        if not zs:
            return np.inf, -np.inf, np.inf, -np.inf
        z, e = np.concatenate(zs), np.concatenate(es)
        return z.min(), z.max(), e.min(), e.max()

    def _compute_axes_limits(self):
        zs, es = self._bounded_points()
        if not zs:
            return -20.0, 20.0, self.E0 - 2.0, self.E0 + 2.0
        z, e = np.concatenate(zs), np.concatenate(es)
        dz, de = z.max() - z.min(), e.max() - e.min()
        px = 0.10 * dz if dz > 0 else 1.0
        py = 0.10 * de if de > 0 else 0.1
        return z.min() - px, z.max() + px, e.min() - py, e.max() + py

    # -- Hamiltonian on a grid (src/FFAG_bucket.py:235-281) --
    @staticmethod
    def cumtrapz_from_grid_start(y, x):
        acc = np.zeros_like(x)
        acc[1:] = np.cumsum(0.5 * (y[:-1] + y[1:]) * np.diff(x))
        return acc

    def A_of_p(self, p_mev):
        return -(self.Tc - np.interp(self.E0 + p_mev, self.Ek_grid, self.T_grid)) / self.Tc

    def B_of_q(self, q_s):
        return self.U0 * (np.sin(self.fi0 + self.w0 * q_s) - np.sin(self.fi0)) / self.Tc / 1e6

    def build_H(self, Nq=241, Np=241):
        """H = Hq(q) + Hp(p) - H(0, 0) on the plot box, Hq = -int B dq and
        Hp = int A dp by the trapezoid rule from the box corner; H_sep at the
        unstable fixed point (zd_u, 0)."""
        self.x_ns = np.linspace(self.x_min_plot, self.x_max_plot, Nq)
        self.y_Ek = np.linspace(self.y_min_plot, self.y_max_plot, Np)
        self.q_grid = self.x_ns * 1e-9
        self.p_grid = self.y_Ek - self.E0
        self.Hp_rel = self.cumtrapz_from_grid_start(self.A_of_p(self.p_grid), self.p_grid)
        self.Hq_rel = -self.cumtrapz_from_grid_start(self.B_of_q(self.q_grid), self.q_grid)
        hq0 = np.interp(0.0, self.q_grid, self.Hq_rel, left=np.nan, right=np.nan)
        hp0 = np.interp(0.0, self.p_grid, self.Hp_rel, left=np.nan, right=np.nan)
        if np.isnan(hq0):   # zero outside the box: the nearer end
            hq0 = self.Hq_rel[0] if abs(self.q_grid[0]) < abs(self.q_grid[-1]) else self.Hq_rel[-1]
        if np.isnan(hp0):
            hp0 = self.Hp_rel[0] if abs(self.p_grid[0]) < abs(self.p_grid[-1]) else self.Hp_rel[-1]
        self.H00 = hq0 + hp0
        self.H_map = self.Hq_rel[:, None] + self.Hp_rel[None, :] - self.H00
        self.H_sep = (np.interp(self.zd_u, self.q_grid, self.Hq_rel) + np.interp(0.0, self.p_grid, self.Hp_rel)
                      - self.H00)

    def H_at(self, q_s, p_mev):
        """H of points (grid interpolation, held at the box edges)."""
        return (np.interp(q_s, self.q_grid, self.Hq_rel) + np.interp(p_mev, self.p_grid, self.Hp_rel) - self.H00)

    # -- bunch length to H threshold and sampling (src/FFAG_bucket.py:325-455) --
    def compute_center_H_vs_span(self):
        """(H, projected length in ns) of the bounded centre-line
        trajectories, sorted by H; the shorter of the forward and backward
        spans when both are bounded."""
        n_sep = self.X_sep.shape[0]
        H_c = self.H_at(self.X_center[:, 0], self.X_center[:, 1])
        H_list, L_list = [], []
        for k in range(self.X_center.shape[0]):
            pid = n_sep + k
            spans = [float(np.ptp(tr[:, pid, 1] * 1e9))
                     for tr, ok in ((self.tr_fwd, self.stable_fwd[pid]), (self.tr_rev, self.stable_rev[pid])) if ok]
            if spans:
                H_list.append(float(H_c[k]))
                L_list.append(min(spans))
        if not H_list:
            raise RuntimeError("no bounded centre-line trajectory to measure the bunch length")
        H_arr, L_arr = np.array(H_list), np.array(L_list)
        o = np.argsort(H_arr)
        self._H_arr_center, self._L_mono_center = H_arr[o], L_arr[o]
        return self._H_arr_center, self._L_mono_center

    def scatter_filter_by_span(self, N_demo, L_target_ns, show=True, rng=None):
        """N_demo points (arrival time [ns], Ek [MeV]) uniform in the sampling
        box with H >= H_thr, H_thr the centre-line H whose trajectory spans
        L_target_ns (clipped to the measured spans, linear interpolation over
        the spans in increasing order); returns (t_ns, Ek_MeV, H_thr)."""
        if self.H_map is None:
            self.build_H()
        if self._H_arr_center is None:
            self.compute_center_H_vs_span()
        H_arr, L_arr = self._H_arr_center, self._L_mono_center
        L_use = float(np.clip(L_target_ns, L_arr.min(), L_arr.max()))
        o = np.argsort(L_arr)
        H_thr = float(_linear_held(L_arr[o], H_arr[o], L_use))
        rng = np.random.default_rng() if rng is None else rng
        acc = np.zeros((0, 2))
        while acc.shape[0] < N_demo:
            x = rng.uniform(self.x_min_sample, self.x_max_sample, size=int(N_demo) * 10)
            y = rng.uniform(self.y_min_sample, self.y_max_sample, size=int(N_demo) * 10)
            keep = self.H_at(x * 1e-9, y - self.E0) >= H_thr
            acc = np.vstack((acc, np.column_stack((x[keep], y[keep]))))
        acc = acc[:N_demo]
        return acc[:, 0], acc[:, 1], H_thr
