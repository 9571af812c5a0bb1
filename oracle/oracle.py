"""FP64 CPU oracle for the tracking hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg use this module, and only as the checker / reported CPU baseline.  The
product package (``pyffa_amd``) never imports it.

It drives the C restatement in ``ffa_oracle.c`` with the reference's control
flow, on the reference's (N, 16) ``LocalBunch`` matrix:

* ``boris_step``  -- ``FFAG_RungeKutta.Boris_solve``, src/FFAG_track.py:948-1049
* ``rk4_step``    -- ``rk4_step_malloc`` + ``RK4_Post_Step_njit``,
  src/FFAG_track.py:51-61, 1743-1801; src/FFAG_MathTools.py:2269-2299
* ``sc_solve``    -- ``FFAG_Bunch.Build_2_5D_SliceGrid_sbs``,
  src/FFAG_Bunch.py:1157-1352 (FFTs in numpy complex128 instead of FFTW
  complex64; grid deposited both the reference way, float32 in particle
  order, and in the fixed-point form the GPU uses)
* ``Tracker.run`` -- the loop of ``rk4_solve_dt_bunch3_withSC_boris`` /
  ``rk4_solve_dt_bunch3_withSC``, src/FFAG_track.py:297-579 / 65-294
  (injection, SC endpoint roll and extrapolation, push, stop condition).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libffa_oracle.so")

Q = 1.60217662e-19
C = 2.99792458e8
E0_J = 938.2723e6 * Q
M0 = E0_J / C ** 2
E0_MEV_BUNCH = 938.2723   # src/FFAG_Bunch.py:90

_lib = None


def build(force: bool = False) -> str:
    """Compile ffa_oracle.c with gcc (OpenMP, no FP contraction)."""
    src = os.path.join(HERE, "ffa_oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", HERE])
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.ora_sc_local_coords.restype = ctypes.c_int64
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


_i64 = ctypes.c_int64
_d = ctypes.c_double
_i = ctypes.c_int


# ---------------------------------------------------------------------------
# Machine description (what GlobalParameters holds in the reference)
# ---------------------------------------------------------------------------
@dataclass
class Machine:
    # Bmap (FFAG_Bfield_analytical attributes, src/FFAG_Field.py:641-706)
    r_min: float
    r_step: float
    r_size: int
    fi_min: float
    fi_step: float
    fi_size: int
    Bz_stack: np.ndarray
    Br_stack: np.ndarray
    Bfi_stack: np.ndarray
    max_order: int
    s_poly: np.ndarray
    sp_poly: np.ndarray
    B0: float
    k_value: float
    nsectors: int
    convex: np.ndarray
    # Emap (FFAG_EField_new attributes, src/FFAG_Field.py:168-198)
    emap_enable: bool = False
    gap_azimuths: np.ndarray = field(default_factory=lambda: np.zeros(0))
    gap_width: float = 1.0
    e_rmin: float = 0.0
    e_rmax: float = 0.0
    rf_shift: np.ndarray = field(default_factory=lambda: np.zeros(1))
    freq_curve: np.ndarray | None = None
    # harmonics (FFAG_BField_Error, src/FFAG_Field.py:965-998)
    harm_enable: bool = False
    harm_m: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    harm_amp: np.ndarray = field(default_factory=lambda: np.zeros(0))
    harm_phase: np.ndarray = field(default_factory=lambda: np.zeros(0))
    harm_rmin: np.ndarray = field(default_factory=lambda: np.zeros(0))
    harm_rmax: np.ndarray = field(default_factory=lambda: np.zeros(0))
    aperture: np.ndarray | None = None
    # spiral gaps (FFAG_EField_spiral attributes, src/FFAG_Field.py:379-430):
    # gap_model "spiral" with maps (nr, nphi) and their axes
    gap_model: str = "fixed"
    emap_er: np.ndarray | None = None
    emap_ez: np.ndarray | None = None
    emap_ephi: np.ndarray | None = None
    emap_r_axis: np.ndarray | None = None
    emap_phi_axis: np.ndarray | None = None

    def gap_regions(self):
        """compute_acc_region + per-gap bounding boxes,
        src/FFAG_Field.py:182-192, 236-261."""
        out = []
        if self.gap_model == "spiral":
            return out
        for th in np.atleast_1d(self.gap_azimuths):
            cs = np.array([self.e_rmin * np.cos(th), self.e_rmin * np.sin(th)])
            ce = np.array([self.e_rmax * np.cos(th), self.e_rmax * np.sin(th)])
            nn = np.array([np.cos(th + np.pi / 2), np.sin(th + np.pi / 2)])
            reg = np.stack([cs + 0.5 * self.gap_width * nn, ce + 0.5 * self.gap_width * nn,
                            ce - 0.5 * self.gap_width * nn, cs - 0.5 * self.gap_width * nn])
            box = (reg[:, 0].min(), reg[:, 0].max(), reg[:, 1].min(), reg[:, 1].max())
            out.append((np.ascontiguousarray(reg[:, 0]), np.ascontiguousarray(reg[:, 1]), box, th))
        return out

    def freq_at(self, t):
        """Interpolation1D_freq_curve -> linear_interpolation_uniform_njit_two_vars,
        src/FFAG_Field.py:309-314, src/FFAG_MathTools.py:1431-1478."""
        fc = self.freq_curve
        if fc is None:
            return 0.0, 0.0
        X0 = fc[0, 0]
        dX = fc[1, 0] - fc[0, 0]
        Y, Z = fc[:, 1], fc[:, 2]
        n = Y.shape[0]
        i = int(np.floor((t - X0) / dX))
        if i < 0:
            i = 0
        elif i >= n - 1:
            i = n - 2
        x0 = X0 + i * dX
        x1 = x0 + dX
        yi = Y[i] + (t - x0) / (x1 - x0) * (Y[i + 1] - Y[i])
        zi = Z[i] + (t - x0) / (x1 - x0) * (Z[i + 1] - Z[i])
        return yi, zi


class Scratch:
    def __init__(self, n):
        self.n = n
        self.fi_vect = np.zeros(n)
        self.Bz = np.zeros(n)
        self.Br = np.zeros(n)
        self.Bfi = np.zeros(n)
        self.Er = np.zeros(n)
        self.Ez = np.zeros(n)
        self.Efi = np.zeros(n)
        self.gap = np.zeros(n, np.int32)
        self.cart = np.zeros(12 * n)
        self.out = np.zeros(6 * n)
        self.dxdt = np.zeros((n, 16))


def field_eval(mc: Machine, M, sc: Scratch, t, use_raw_fi_for_emap=True):
    """Field part shared by Boris_solve and FunctionForAccelerationBunch_dt:
    fast_mod, GetBfields, AddHigherOrderBmaps, harmonics, Emap gaps.
    Returns (f_rf, v0) sampled at t."""
    L = lib()
    n = M.shape[0]
    L.ora_fast_mod(_p(M), _i64(n), _d(np.pi * 2 / mc.nsectors), _p(sc.fi_vect))
    L.ora_get_bfields(_p(M), _i64(n), _p(sc.fi_vect), _p(sc.Bz), _p(sc.Br), _p(sc.Bfi),
                      _p(mc.s_poly), _i(mc.s_poly.size), _p(mc.sp_poly), _i(mc.sp_poly.size),
                      _d(mc.B0), _d(mc.k_value), _i(mc.nsectors), _p(mc.convex), _i(mc.convex.shape[0]))
    if mc.max_order >= 1:
        L.ora_bmap_add_all_orders(_d(mc.r_min), _d(mc.r_step), _i64(mc.r_size), _d(mc.fi_min), _d(mc.fi_step),
                                  _i64(mc.fi_size), _p(mc.Bz_stack), _p(mc.Bfi_stack), _p(mc.Br_stack),
                                  _p(M), _i64(n), _p(sc.fi_vect), _i(mc.max_order),
                                  _p(sc.Bz), _p(sc.Bfi), _p(sc.Br))
    if mc.harm_enable:
        L.ora_bharmonics(_p(M), _i64(n), _i(mc.harm_m.size), _p(mc.harm_m), _p(mc.harm_amp), _p(mc.harm_phase),
                         _p(mc.harm_rmin), _p(mc.harm_rmax), _p(sc.Bz), _p(sc.Br), _p(sc.Bfi))
    f_rf, v0 = mc.freq_at(t)
    if not mc.emap_enable:
        sc.Er.fill(0.0)
        sc.Efi.fill(0.0)
        sc.Ez.fill(0.0)
    if mc.gap_model == "spiral":
        if mc.emap_enable:
            ra, pa = mc.emap_r_axis, mc.emap_phi_axis
            gaz = np.ascontiguousarray(np.atleast_1d(mc.gap_azimuths), dtype=np.float64)
            L.ora_emap_spiral(_p(M), _i64(n), _p(gaz), _i(gaz.size), _d(v0 * (1.0 / mc.gap_width)),
                              _d(ra[0]), _d(ra[1] - ra[0]), _i64(ra.size), _d(pa[0]), _d(pa[1] - pa[0]),
                              _i64(pa.size), _p(mc.emap_er), _p(mc.emap_ez), _p(mc.emap_ephi),
                              _p(sc.Er), _p(sc.Ez), _p(sc.Efi), _p(sc.gap))
        else:
            sc.gap.fill(0)
        return f_rf, v0
    for g, (px, py, box, th) in enumerate(mc.gap_regions()):
        L.ora_emap_gap(_p(M), _i64(n), _d((1.0 / mc.gap_width) * v0), _d(th + np.pi / 2.0), _p(px), _p(py),
                       _i(px.size), _d(box[0]), _d(box[1]), _d(box[2]), _d(box[3]),
                       _p(sc.Er), _p(sc.Efi), _p(sc.Ez), _p(sc.gap), _i(g))
    return f_rf, v0


def boris_step(mc: Machine, M, sc: Scratch, t, dt):
    """One Boris_solve call on LocalBunch M (modified in place)."""
    L = lib()
    n = M.shape[0]
    f_rf, v0 = field_eval(mc, M, sc, t)
    dRF = f_rf * np.pi * 2 * dt
    rf_shift = np.ascontiguousarray(mc.rf_shift, dtype=np.float64)
    L.ora_cyl2cart(_p(M), _i64(n), _p(sc.Er), _p(sc.Ez), _p(sc.Efi), _p(sc.Br), _p(sc.Bz), _p(sc.Bfi),
                   _p(sc.gap), _p(rf_shift), _i(1 if mc.emap_enable else 0), _p(sc.cart))
    ap = mc.aperture if mc.aperture is not None else np.zeros(4)
    L.ora_boris(_p(M), _i64(n), _p(sc.cart), _d(dt), _d(Q), _d(M0), _d(C),
                _i(1 if mc.aperture is not None else 0), _p(np.ascontiguousarray(ap, dtype=np.float64)), _p(sc.out))
    L.ora_cart2cyl(_p(M), _i64(n), _p(sc.out), _p(sc.cart), _d(dRF))


def _rk4_func(mc: Machine, t, S, sc: Scratch):
    """FunctionForAccelerationBunch_dt, src/FFAG_track.py:1743-1801."""
    L = lib()
    n = S.shape[0]
    f_rf, v0 = field_eval(mc, S, sc, t)
    rate = f_rf * np.pi * 2
    rf_shift = np.ascontiguousarray(mc.rf_shift, dtype=np.float64)
    L.ora_rk4_rhs(_p(S), _i64(n), _p(sc.Er), _p(sc.Ez), _p(sc.Efi), _p(sc.Br), _p(sc.Bz), _p(sc.Bfi),
                  _d(Q), _d(C), _d(E0_J), _p(sc.dxdt), _d(rate), _p(rf_shift), _p(sc.gap),
                  _i(1 if mc.emap_enable else 0))
    return sc.dxdt


def rk4_step(mc: Machine, M, sc: Scratch, t, h):
    """rk4_step_malloc then RK4_Post_Step_njit (state column 5 = E_total [J])."""
    k1 = h * _rk4_func(mc, t, M, sc)
    k2 = h * _rk4_func(mc, t + 0.5 * h, np.ascontiguousarray(M + 0.5 * k1), sc)
    k3 = h * _rk4_func(mc, t + 0.5 * h, np.ascontiguousarray(M + 0.5 * k2), sc)
    k4 = h * _rk4_func(mc, t + h, np.ascontiguousarray(M + k3), sc)
    d = np.ascontiguousarray((k1 + 2 * k2 + 2 * k3 + k4) / 6)
    lib().ora_rk4_post(_p(M), _i64(M.shape[0]), _p(d))


# ---------------------------------------------------------------------------
# 2.5D space charge
# ---------------------------------------------------------------------------
@dataclass
class SEO:
    Ek_axis: np.ndarray
    fi_axis: np.ndarray
    r_matrix: np.ndarray
    pr_matrix: np.ndarray
    perimeter: np.ndarray


def _pymod(a, b):
    return a % b     # Python float modulo (sign of the divisor)


def fixed_point_bits(n_particles: int) -> int:
    """Fractional bits per axis of the fixed-point CIC weights: at most 20 and
    small enough that n_particles * 2^(2 bits) < 2^63 (no int64 overflow even
    if every particle lands in one cell)."""
    b = 20
    while b > 1 and n_particles * (1 << (2 * b)) >= (1 << 63):
        b -= 1
    return b


def sc_solve(M, step_survive, seo: SEO, ns, nx, nz, sigma_ratio, marcosize, rk=False,
             bounds_override=None, want_ref_rho=False):
    """Build_2_5D_SliceGrid_sbs on LocalBunch M: writes Esc into columns 10-12
    of the survivors and returns a dict of intermediate products."""
    L = lib()
    n = M.shape[0]
    loc = np.zeros(6 * n)
    sg = ctypes.c_double(0.0)
    cnt = L.ora_sc_local_coords(_p(M), _i64(n), _p(step_survive.astype(np.uint8)), _d(C), _i(1 if rk else 0),
                                _p(loc), ctypes.byref(sg))
    r, fi, z, gam, fim, idx = (loc[j * n: j * n + cnt].copy() for j in range(6))
    arc = np.zeros(3)
    L.ora_divide_arc2(_p(np.ascontiguousarray(fim)), _i64(cnt), _p(arc))
    s_, e_ = arc[0], arc[1]
    if e_ < s_:
        e_ += 2 * np.pi
    fi_s, fi_e = _pymod(s_, 2 * np.pi), _pymod(e_, 2 * np.pi)
    if fi_e <= fi_s:
        fi_e += 2 * np.pi
    step = (fi_e - fi_s) / (ns - 1)
    fi_grid = np.linspace(fi_s, fi_e, ns)
    mean_gm1 = sg.value / cnt - 1.0 if cnt > 0 else 0.0
    ek = np.ascontiguousarray((gam - 1) * E0_MEV_BUNCH)
    r0 = np.zeros(cnt)
    pr0 = np.zeros(cnt)
    ax = seo.Ek_axis
    fx = seo.fi_axis
    for tab, out in ((seo.r_matrix, r0), (seo.pr_matrix, pr0)):
        L.ora_bilinear_uniform(_d(ax[0]), _d(ax[1] - ax[0]), _i64(ax.size), _d(fx[0]), _d(fx[1] - fx[0]),
                               _i64(fx.size), _p(np.ascontiguousarray(tab, dtype=np.float64)),
                               _p(ek), _p(np.ascontiguousarray(fim)), _i64(cnt), _p(out))
    x = np.zeros(cnt)
    sl = np.zeros(cnt, np.int32)
    counts = np.zeros(ns, np.int32)
    sums = np.zeros(4)
    L.ora_sc_slices(_p(r), _p(z), _p(np.ascontiguousarray(fim)), _p(r0), _i64(cnt), _d(step), _d(fi_s), _i(ns),
                    _p(x), _p(sl), _p(counts), _p(sums))
    if cnt > 1:
        mx, mz = sums[0] / cnt, sums[1] / cnt
        vx = sums[2] / cnt - mx ** 2
        vz = sums[3] / cnt - mz ** 2
        sx, sz = np.sqrt(max(vx, 0.0)), np.sqrt(max(vz, 0.0))
    else:
        mx = mz = sx = sz = 0.0
    perim = np.interp(mean_gm1 * E0_MEV_BUNCH, seo.Ek_axis, seo.perimeter)
    bl = perim * (fi_e - fi_s) / np.pi / 2.0
    s_grid = fi_grid / np.pi / 2.0 * perim
    xmin, xmax = mx - sigma_ratio * sx, mx + sigma_ratio * sx
    zmin, zmax = mz - sigma_ratio * sz, mz + sigma_ratio * sz
    if bounds_override is not None:
        xmin, xmax, zmin, zmax = bounds_override
    q_part = marcosize * Q
    dx = (xmax - xmin) / nx
    dz = (zmax - zmin) / nz
    order = np.argsort(sl, kind="stable").astype(np.int64)
    res = dict(count=cnt, fi_s=fi_s, fi_e=fi_e, slice_step=step, mean_gamma_m1=mean_gm1, sums=sums,
               counts=counts, x=x, z=z, slice_idx=sl, r0=r0, pr0=pr0, idx=idx,
               xmin=xmin, xmax=xmax, zmin=zmin, zmax=zmax, bunch_length=bl, s_grid=s_grid, dx=dx, dz=dz)
    bits = fixed_point_bits(max(cnt, 1))
    grid = np.zeros((ns, nz, nx), np.int64)
    cell_count = np.zeros((ns, nz, nx), np.int32)
    L.ora_deposit_fixed(_p(x), _p(z), _i64(cnt), _p(sl), _d(xmin), _d(xmax), _d(zmin), _d(zmax),
                        _i(nx), _i(nz), _i(ns), _i(bits), _p(grid), _p(cell_count))
    res.update(grid_fixed=grid, cell_count=cell_count, bits=bits)
    if want_ref_rho:
        rho32 = np.zeros((ns, nz, nx), np.float32)
        L.ora_deposit_f32(_p(x), _p(z), _p(order), _i64(cnt), _p(sl), _d(q_part), _d(bl), _d(xmin), _d(xmax),
                          _d(zmin), _d(zmax), _i(nx), _i(nz), _i(ns), _p(rho32))
        res["rho_ref_f32"] = rho32
    if dx <= 0.0 or dz <= 0.0 or bl <= 0.0 or cnt == 0:
        res.update(bad_grid=True)
        return res
    qscale = (q_part / bl * ns) * (1.0 / (dx * dz))
    rho = grid.astype(np.float64) * (qscale / float(1 << (2 * bits)))
    G = np.zeros((nz, nx), np.float32)
    L.ora_green(_i(nx), _i(nz), _d(dx), _d(dz), _p(G))
    phi = np.real(np.fft.ifft2(np.fft.fft2(rho, axes=(1, 2)) * G.astype(np.float64)[None], axes=(1, 2)))
    phi = np.ascontiguousarray(phi)
    Ex = np.zeros_like(phi)
    Ez = np.zeros_like(phi)
    for k in range(ns):
        L.ora_field_from_potential(_p(phi[k]), _i(nz), _i(nx), _d(dx), _d(dz), _d(bl), _d(mean_gm1),
                                   _p(Ex[k]), _p(Ez[k]))
    Ef = np.zeros(ns)
    L.ora_total_field_1d(_p(np.ascontiguousarray(s_grid)), _p(counts.astype(np.float64) * q_part), _i(ns), _p(Ef))
    L.ora_sc_gather(_p(Ex), _p(Ez), _i(ns), _i(nz), _i(nx), _p(x), _p(z), _p(pr0), _p(sl), _p(idx), _i64(cnt),
                    _d(xmin), _d(zmin), _d(dx), _d(dz), _d(bl), _p(M))
    res.update(bad_grid=False, rho=rho, phi=phi, Ex=Ex, Ez=Ez, Ef=Ef, green=G, qscale=qscale)
    return res


# ---------------------------------------------------------------------------
# The tracking loop
# ---------------------------------------------------------------------------
class Tracker:
    """Restatement of the reference loop on a single rank.  ``M`` is the
    (N, 16) LocalBunch sorted by injection time."""

    def __init__(self, mc: Machine, M, *, solver="boris", enable_sc=False, sc_step=20,
                 grid=(64, 64, 64, 4.0), marcosize=1.0, seo: SEO | None = None):
        self.mc = mc
        self.M = np.ascontiguousarray(M, dtype=np.float64)
        n = self.M.shape[0]
        self.static = self.M.copy()
        self.sc = Scratch(n)
        self.solver = solver
        self.enable_sc = enable_sc
        self.sc_step = int(sc_step)
        self.ns, self.nx, self.nz, self.sigma_ratio = int(grid[0]), int(grid[1]), int(grid[2]), float(grid[3])
        self.marcosize = marcosize
        self.seo = seo
        self.head = 0
        self.step_survive = np.zeros(n, bool)
        self.E_pre = np.zeros((n, 3))
        self.E_prepre = np.zeros((n, 3))
        self.m_pre = np.zeros(n, bool)
        self.m_prepre = np.zeros(n, bool)
        self.E_p = np.zeros((n, 3))
        self.injected = 0
        self.last_sc = None

    def inject(self, t, dt):
        """InjectParticles, src/FFAG_Bunch.py:265-303 (inj_t in ns)."""
        hi = t * 1e9
        lo = (t - dt) * 1e9
        tail = self.static[self.head:, 6]
        a = np.searchsorted(tail, lo, side="left")
        b = np.searchsorted(tail, hi, side="right")
        if b - a > 0:
            s, e = self.head + a, self.head + b
            self.M[s:e, 7] = 1
            self.M[s:e, 8] = 1
            self.step_survive[s:e] = True
            self.head = e
            self.injected += b - a

    def run(self, t_start, dt, nsteps, max_turn=None, max_time=None, start_step=0):
        """Steps start_step .. nsteps-1 (start_step > 0 continues a run)."""
        t = t_start
        steps_i = start_step
        M = self.M
        while steps_i < nsteps:
            self.inject(t, dt)
            self.step_survive[:] = M[:, 8] != 0          # UpdataPreStepNjit
            if self.enable_sc and steps_i % self.sc_step == 0:
                self.last_sc = sc_solve(M, self.step_survive, self.seo, self.ns, self.nx, self.nz,
                                        self.sigma_ratio, self.marcosize, rk=(self.solver == "rk4"))
                alive = M[:, 8] != 0
                if steps_i == 0:
                    self.m_pre[:] = alive
                    self.m_prepre[:] = False
                    self.E_pre[alive] = M[alive, 10:13]
                    self.E_prepre[:] = self.E_pre
                else:
                    self.m_prepre[:] = self.m_pre
                    self.m_pre[:] = alive
                    self.E_prepre[:] = self.E_pre
                    self.E_pre[alive] = M[alive, 10:13]
            k = steps_i % self.sc_step
            s = k / self.sc_step if steps_i >= self.sc_step else 0.0
            both = self.m_pre & self.m_prepre
            only = self.m_pre & ~self.m_prepre
            self.E_p[both] = self.E_pre[both] + s * (self.E_pre[both] - self.E_prepre[both])
            self.E_p[only] = self.E_pre[only]
            M[self.m_pre, 10:13] = self.E_p[self.m_pre]
            if self.solver == "boris":
                boris_step(self.mc, M, self.sc, t, dt)
            else:
                rk4_step(self.mc, M, self.sc, t, dt)
            t = t + dt
            steps_i += 1
            if steps_i % 1000 == 0:
                fi = M[:, 4]
                fi = fi[~np.isnan(fi)]
                if max_turn is not None and abs(np.mean(fi)) > max_turn * np.pi * 2:
                    break
                if max_time is not None and t > max_time:
                    break
        self.t = t
        self.steps = steps_i
        return steps_i


# ---------------------------------------------------------------------------
# C multi-step driver (no space charge)
# ---------------------------------------------------------------------------
class _OraMachine(ctypes.Structure):
    _fields_ = [("r_min", _d), ("r_step", _d), ("fi_min", _d), ("fi_step", _d), ("r_size", _i64), ("fi_size", _i64),
                ("Bz_stack", ctypes.c_void_p), ("Br_stack", ctypes.c_void_p), ("Bfi_stack", ctypes.c_void_p),
                ("max_order", ctypes.c_int32), ("s_poly", ctypes.c_void_p), ("n_s", ctypes.c_int32),
                ("sp_poly", ctypes.c_void_p), ("n_sp", ctypes.c_int32), ("B0", _d), ("k_value", _d),
                ("nsectors", ctypes.c_int32), ("convex", ctypes.c_void_p), ("m_seg", ctypes.c_int32),
                ("n_harm", ctypes.c_int32), ("harm_m", ctypes.c_void_p), ("harm_amp", ctypes.c_void_p),
                ("harm_phase", ctypes.c_void_p), ("harm_rmin", ctypes.c_void_p), ("harm_rmax", ctypes.c_void_p),
                ("emap_enable", ctypes.c_int32), ("n_gap", ctypes.c_int32), ("gap_width", _d),
                ("gap_theta", ctypes.c_void_p), ("gap_px", ctypes.c_void_p), ("gap_py", ctypes.c_void_p),
                ("gap_box", ctypes.c_void_p), ("rf_shift", ctypes.c_void_p), ("freq_curve", ctypes.c_void_p),
                ("n_freq", _i64), ("freq_cols", ctypes.c_int32), ("aperture_enable", ctypes.c_int32),
                ("aperture", ctypes.c_void_p), ("q", _d), ("c", _d), ("m0", _d), ("E0", _d),
                ("gap_model", ctypes.c_int32), ("emap_nr", _i64), ("emap_nphi", _i64), ("emap_r0", _d),
                ("emap_dr", _d), ("emap_phi0", _d), ("emap_dphi", _d), ("emap_er", ctypes.c_void_p),
                ("emap_ez", ctypes.c_void_p), ("emap_ephi", ctypes.c_void_p)]


def _c_machine(mc: Machine):
    keep = []

    def arr(a, dt=np.float64):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a.ctypes.data

    m = _OraMachine()
    m.r_min, m.r_step, m.fi_min, m.fi_step = mc.r_min, mc.r_step, mc.fi_min, mc.fi_step
    m.r_size, m.fi_size = mc.r_size, mc.fi_size
    m.Bz_stack, m.Br_stack, m.Bfi_stack = arr(mc.Bz_stack), arr(mc.Br_stack), arr(mc.Bfi_stack)
    m.max_order = mc.max_order
    m.s_poly, m.n_s = arr(mc.s_poly), mc.s_poly.size
    m.sp_poly, m.n_sp = arr(mc.sp_poly), mc.sp_poly.size
    m.B0, m.k_value, m.nsectors = mc.B0, mc.k_value, mc.nsectors
    m.convex, m.m_seg = arr(mc.convex), mc.convex.shape[0]
    if mc.harm_enable:
        m.n_harm = mc.harm_m.size
        m.harm_m = arr(mc.harm_m, np.int64)
        m.harm_amp, m.harm_phase = arr(mc.harm_amp), arr(mc.harm_phase)
        m.harm_rmin, m.harm_rmax = arr(mc.harm_rmin), arr(mc.harm_rmax)
    regs = mc.gap_regions()
    m.emap_enable = 1 if mc.emap_enable else 0
    m.n_gap = len(regs)
    m.gap_width = mc.gap_width
    if regs:
        m.gap_theta = arr([r[3] for r in regs])
        m.gap_px = arr(np.stack([r[0] for r in regs]))
        m.gap_py = arr(np.stack([r[1] for r in regs]))
        m.gap_box = arr(np.array([r[2] for r in regs]))
    m.rf_shift = arr(mc.rf_shift)
    if mc.freq_curve is not None:
        fc = np.ascontiguousarray(mc.freq_curve, dtype=np.float64)
        m.freq_curve, m.n_freq, m.freq_cols = arr(fc), fc.shape[0], fc.shape[1]
    if mc.aperture is not None:
        m.aperture_enable = 1
        m.aperture = arr(mc.aperture)
    m.q, m.c, m.m0, m.E0 = Q, C, M0, E0_J
    if mc.gap_model == "spiral":
        gaz = np.atleast_1d(mc.gap_azimuths)
        m.gap_model = 1
        m.n_gap = gaz.size
        m.gap_theta = arr(gaz)
        ra, pa = mc.emap_r_axis, mc.emap_phi_axis
        m.emap_nr, m.emap_nphi = ra.size, pa.size
        m.emap_r0, m.emap_dr, m.emap_phi0, m.emap_dphi = ra[0], ra[1] - ra[0], pa[0], pa[1] - pa[0]
        m.emap_er, m.emap_ez, m.emap_ephi = arr(mc.emap_er), arr(mc.emap_ez), arr(mc.emap_ephi)
    return m, keep


def run_c(mc: Machine, M, t0, dt, nsteps, rk4=False, head=0):
    """Run nsteps Boris / RK4 steps (no SC) on M in place with the C driver;
    injection follows M's inj_t column (sorted).  Returns (t_end, head)."""
    L = lib()
    L.ora_run.restype = ctypes.c_int64
    m, keep = _c_machine(mc)
    inj_t = np.ascontiguousarray(M[:, 6].copy())
    h = ctypes.c_int64(head)
    rc = L.ora_run(ctypes.byref(m), _p(M), _i64(M.shape[0]), _p(inj_t), ctypes.byref(h), _d(t0), _d(dt),
                   _i64(nsteps), _i(1 if rk4 else 0))
    if rc < 0:
        raise MemoryError("oracle scratch allocation failed")
    t = t0
    for _ in range(nsteps):
        t = t + dt
    return t, h.value
