"""Field objects of the tracking path (host side).

Same constructors and attributes as the reference's ``src/FFAG_Field.py``:

* ``FFAG_Bfield_analytical(config_Bmap_path, max_order)`` (:635-944) -- the
  analytic order-0 scaling field plus the Taylor-coefficient stacks
  ``Bz/Br/Bfi_order_<n>.npz``;
* ``FFAG_EField_new(freq_curve_path, EnableFlag)`` (:167-376) -- fixed RF gaps
  and the frequency/voltage curve;
* ``FFAG_BField_Error(harmonics_cfg, EnableFlag)`` (:947-1042) -- harmonic
  field errors.

The per-particle evaluation of these fields happens inside the HIP kernels;
these classes hold the parameters and build the device-resident tables
(``device_taylor``) and the C-ABI machine description (``build_machine``).
"""
from __future__ import annotations

import ast
import json
import math
import os

import numpy as np

from . import _lib


class FFAG_Bfield_analytical:
    """Analytic order-0 field + Taylor stacks (src/FFAG_Field.py:635-706)."""

    def __init__(self, config_Bmap_path=None, max_order=0, flag3D=True, AddNumricalMap=True, *, _cfg=None,
                 _stacks=None):
        if _cfg is None:
            with open(config_Bmap_path, "r") as f:
                _cfg = json.load(f)
        bm = _cfg["Bmap"]
        rmin, rmax, rstep = bm["rmin_max_step_m"]
        self.Nsectors = int(bm["NSector"])
        theta_step = bm["theta_step_rad"]
        n_fi = int((2 * np.pi / self.Nsectors) / theta_step) + 1
        fi_axis = np.linspace(0, 2 * np.pi / self.Nsectors, n_fi)
        n_r = int((rmax - rmin) / rstep) + 1
        r_axis = np.linspace(rmin, rmax, n_r)
        s_val = self.calculate_s(r_axis, 0.0, bm["SpiralAngle_deg"])
        self.s_rad_polynomial = np.polyfit(r_axis, s_val, 6)
        n = len(self.s_rad_polynomial) - 1
        self.s_rad_polynomial_prime = np.array([self.s_rad_polynomial[i] * (n - i) for i in range(n)])
        self.B0, self.k_value = bm["coefficients_tupel"]
        self.r_axis, self.fi_axis = r_axis, fi_axis
        self.r_min, self.r_max, self.r_step = rmin, rmax, rstep
        self.fi_min, self.fi_max = fi_axis[0], fi_axis[-1]
        self.fi_step = fi_axis[1] - fi_axis[0]
        self.r_size, self.fi_size = len(r_axis), len(fi_axis)
        self.convex_params = np.array(self.BmapParams_to_ConvexParams(bm), dtype=np.float64).reshape(-1, 5)
        self.max_order = int(max_order)
        self.flag3D = flag3D
        self.config_path = config_Bmap_path
        self.BmapFoldName = os.path.dirname(config_Bmap_path) if config_Bmap_path else None
        self._dev_cache = {}
        if _stacks is not None:
            self.Bz_stack, self.Br_stack, self.Bfi_stack = (np.ascontiguousarray(s[: self.max_order + 1])
                                                            for s in _stacks)
        elif AddNumricalMap:
            fold = os.path.normpath(self.BmapFoldName)
            bz, br, bfi = [], [], []
            for k in range(self.max_order + 1):
                br.append(np.load(os.path.join(fold, f"Br_order_{k}.npz"))["Br_Taylor_coeff"])
                bz.append(np.load(os.path.join(fold, f"Bz_order_{k}.npz"))["Bz_Taylor_coeff"])
                bfi.append(np.load(os.path.join(fold, f"Bfi_order_{k}.npz"))["Bfi_Taylor_coeff"])
            self.Br_stack = np.ascontiguousarray(np.stack(br))
            self.Bz_stack = np.ascontiguousarray(np.stack(bz))
            self.Bfi_stack = np.ascontiguousarray(np.stack(bfi))
        else:
            self.Bz_stack = self.Br_stack = self.Bfi_stack = None
        if self.Bz_stack is not None and self.Bz_stack.shape[1:] != (self.r_size, self.fi_size):
            raise ValueError(f"Taylor stacks have shape {self.Bz_stack.shape[1:]}, grid is "
                             f"({self.r_size}, {self.fi_size})")

    @classmethod
    def from_arrays(cls, config: dict, max_order: int, stacks):
        """Build from an in-memory config_Bmap dict and (Bz, Br, Bfi) stacks."""
        return cls(None, max_order, _cfg=config, _stacks=stacks)

    @staticmethod
    def BmapParams_to_ConvexParams(Bmap):
        """Segment table (x0, x1, W0, |polarity|, sign) in radians for the
        non-zero polarities (src/FFAG_Field.py:732-767)."""
        rows = []
        sector = 360.0 / Bmap["NSector"]
        a = 0.0
        for frac, pol, fr in zip(Bmap["interval_1"], Bmap["positive_or_negative_1"], Bmap["fringe_width_1"]):
            if pol == 0:
                a += frac * sector
                continue
            b = a + frac * sector
            rows.append((np.deg2rad(a), np.deg2rad(b), np.deg2rad((b - a) * fr), float(abs(pol)),
                         1.0 if pol > 0 else -1.0))
            a = b
        return rows

    @staticmethod
    def calculate_s(r_axis, s0, alpha_deg):
        """Spiral azimuth s(r) by accumulating tan(alpha)/r dr (src/FFAG_Field.py:770-797)."""
        s = np.zeros_like(r_axis)
        s[0] = s0
        ta = np.tan(np.deg2rad(alpha_deg))
        for i in range(1, len(r_axis)):
            s[i] = s[i - 1] + ta / r_axis[i] * (r_axis[i] - r_axis[i - 1])
        return s

    def device_taylor(self, dtype, device):
        """Taylor orders 1..max_order interleaved [phi][r][order][Bz, Bfi, Br]
        on the device (cached per dtype).  Azimuth-major: particles that
        share an azimuth cell (neighbours after the FP32 tier's azimuth
        ordering) read neighbouring r cells, a few cache lines per wave."""
        import torch
        key = (str(dtype), str(device))
        if key in self._dev_cache:
            return self._dev_cache[key]
        if self.max_order < 1:
            t = torch.zeros(1, dtype=dtype, device=device)
        else:
            parts = []
            for k in range(1, self.max_order + 1):
                for st in (self.Bz_stack, self.Bfi_stack, self.Br_stack):
                    parts.append(torch.from_numpy(np.ascontiguousarray(st[k])).to(device=device, dtype=dtype))
            t = torch.stack(parts, dim=-1).permute(1, 0, 2).contiguous()    # (NPHI, NR, 3*max_order)
        self._dev_cache[key] = t
        return t


class FFAG_EField_new:
    """Fixed RF gaps (src/FFAG_Field.py:167-376)."""

    def __init__(self, freq_curve_path, EnableFlag):
        header, data = self.ParseEmapFile(freq_curve_path)
        self.GapType = "fixed"
        self.acc_phi0 = np.deg2rad(header.get("acc_phi_paint", 0.0))
        self.gap_azimuths = np.deg2rad(np.atleast_1d(header["gap_azimuth"]).astype(np.float64))
        self.gap_width = header["gap_width"]
        self.rmin = header["E_rmin"]
        self.rmax = header["E_rmax"]
        self.freq_curve = data
        self.harmonic = header["harmonic"]
        self.EnableFlag = bool(EnableFlag)
        self.rf_shift = np.deg2rad(np.atleast_1d(header["rf_shift"]).astype(np.float64))
        self.acc_regions = self.compute_acc_region()
        self.box_bounds = [(r[:, 0].min(), r[:, 0].max(), r[:, 1].min(), r[:, 1].max()) for r in self.acc_regions]
        self.acc_Etot_uniform = 1 / self.gap_width

    def compute_acc_region(self):
        """Rectangle of half-width gap_width/2 about each gap's radial line."""
        regions = []
        for th in np.atleast_1d(self.gap_azimuths):
            a = np.array([self.rmin * np.cos(th), self.rmin * np.sin(th)])
            b = np.array([self.rmax * np.cos(th), self.rmax * np.sin(th)])
            nrm = np.array([np.cos(th + np.pi / 2), np.sin(th + np.pi / 2)])
            h = 0.5 * self.gap_width * nrm
            regions.append(np.stack([a + h, b + h, b - h, a - h]))
        return regions

    def Interpolation1D_freq_curve(self, time):
        """f_RF(t), V0(t) by clamped linear interpolation on the uniform time
        grid (src/FFAG_Field.py:309-314)."""
        fc = self.freq_curve
        x0t = fc[0, 0]
        dX = fc[1, 0] - fc[0, 0]
        n = fc.shape[0]
        i = int(math.floor((time - x0t) / dX))
        i = 0 if i < 0 else (n - 2 if i >= n - 1 else i)
        xa = x0t + i * dX
        xb = xa + dX
        f = fc[i, 1] + (time - xa) / (xb - xa) * (fc[i + 1, 1] - fc[i, 1])
        v = fc[i, 2] + (time - xa) / (xb - xa) * (fc[i + 1, 2] - fc[i, 2])
        return f, v

    @staticmethod
    def ParseEmapFile(file_path):
        """`key = value` header up to the `#Time (s)` line, then the numeric
        table (src/FFAG_Field.py:316-376)."""
        params = {}
        header_lines = 0
        with open(file_path, "r") as fh:
            for line in fh:
                s = line.strip()
                header_lines += 1
                if s.startswith("#Time (s)"):
                    break
                if "=" not in s:
                    continue
                key, value = (t.strip() for t in s.split("=", 1))
                try:
                    params[key] = float(value) if ("." in value or "e" in value.lower()) else int(value)
                    continue
                except ValueError:
                    pass
                parsed = None
                if (value.startswith("[") and value.endswith("]")) or (value.startswith("(") and value.endswith(")")):
                    try:
                        obj = ast.literal_eval(value)
                        if isinstance(obj, (list, tuple)):
                            parsed = [float(x) for x in obj]
                    except (ValueError, SyntaxError):
                        parsed = None
                if parsed is None and "," in value:
                    try:
                        parsed = [float(p) for p in value.split(",") if p.strip()]
                    except ValueError:
                        parsed = None
                params[key] = parsed if parsed is not None else value
        data = np.loadtxt(file_path, skiprows=header_lines)
        return params, data


class FFAG_EField_spiral:
    """Spiral RF gaps (src/FFAG_Field.py:379-550): the same frequency-curve
    file as FFAG_EField_new with ``gap_model = spiral``, plus three maps
    (``Er_map_file``, ``Ez_map_file``, ``Ephi_map_file``; first row the
    continuous azimuth axis, first column the radius) per unit V0/gap_width.
    The field evaluation runs on the GPU (gap_model 1 of the machine
    description); ``device_emap`` packs the maps for it."""

    def __init__(self, freq_curve_path, EnableFlag):
        header, data = self.ParseEmapFile(freq_curve_path)
        self.GapType = "spiral"
        self.acc_phi0 = np.deg2rad(float(header["acc_phi_paint"]))
        self.gap_azimuths = np.deg2rad(np.atleast_1d(header["gap_azimuth"]).astype(np.float64))
        self.gap_width = float(header["gap_width"])
        self.freq_curve = data
        self.harmonic = int(header["harmonic"])
        self.EnableFlag = bool(EnableFlag)
        self.rf_shift = np.deg2rad(np.atleast_1d(header["rf_shift"]).astype(np.float64))
        d = os.path.dirname(os.path.abspath(freq_curve_path))
        self.Er_map, self.r_axis, self.phi_axis = self.LoadSpiralMap(
            os.path.join(d, header.get("Er_map_file", "Er_coef.txt")))
        self.Ezcoef_map, _, _ = self.LoadSpiralMap(os.path.join(d, header.get("Ez_map_file", "Ez_coef.txt")))
        self.Ephi_map, _, _ = self.LoadSpiralMap(os.path.join(d, header.get("Ephi_map_file", "Ephi_coef.txt")))
        self.r0 = self.r_axis[0]
        self.dr = self.r_axis[1] - self.r_axis[0]
        self.nr = len(self.r_axis)
        self.phi0 = self.phi_axis[0]
        self.dphi = self.phi_axis[1] - self.phi_axis[0]
        self.nphi = len(self.phi_axis)
        self.map_rmin, self.map_rmax = self.r_axis[0], self.r_axis[-1]
        self.map_phimin, self.map_phimax = self.phi_axis[0], self.phi_axis[-1]
        self.acc_Etot_uniform = 1.0 / self.gap_width
        # gap intersections with the closed orbits, by energy
        order = np.argsort(np.atleast_1d(header.get("Ek_gap", [0.0])).astype(np.float64))
        self.Ek_gap = np.atleast_1d(header.get("Ek_gap", [0.0])).astype(np.float64)[order]
        self.r_gap = np.atleast_1d(header.get("r_gap", [0.0])).astype(np.float64)[order]
        self.phi_gap = np.deg2rad(np.atleast_1d(header.get("phi_gap_deg", [0.0])).astype(np.float64))[order]
        self.phi_gap_unwrapped = np.deg2rad(
            np.atleast_1d(header.get("phi_gap_unwrapped_deg", [0.0])).astype(np.float64))[order]
        self._dev_cache = {}

    def r_gap_interp(self, ek):
        return np.interp(np.asarray(ek, dtype=np.float64), self.Ek_gap, self.r_gap)

    def phi_gap_interp_unwrapped(self, ek):
        return np.interp(np.asarray(ek, dtype=np.float64), self.Ek_gap, self.phi_gap_unwrapped)

    @staticmethod
    def LoadSpiralMap(filepath):
        raw = np.loadtxt(filepath)
        return raw[1:, 1:].astype(np.float64), raw[1:, 0].astype(np.float64), raw[0, 1:].astype(np.float64)

    Interpolation1D_freq_curve = FFAG_EField_new.Interpolation1D_freq_curve
    ParseEmapFile = staticmethod(FFAG_EField_new.ParseEmapFile)

    def device_emap(self, device):
        """The three maps interleaved [r][phi][Er, Ez/z, Ephi], float64, on the
        device (cached)."""
        import torch
        key = str(device)
        if key not in self._dev_cache:
            m = np.stack([self.Er_map, self.Ezcoef_map, self.Ephi_map], axis=-1)
            self._dev_cache[key] = torch.from_numpy(np.ascontiguousarray(m)).to(device)
        return self._dev_cache[key]


class FFAG_BField_Error:
    """Harmonic Bz errors (src/FFAG_Field.py:947-1042)."""

    def __init__(self, harmonics_cfg, EnableFlag):
        self.harmonics_cfg = harmonics_cfg
        self.enable_flag = bool(EnableFlag)
        self.Nharm = len(harmonics_cfg)
        self.m_arr = np.array([int(h["m"]) for h in harmonics_cfg], dtype=np.int64)
        self.amp_arr = np.array([float(h["amp_Gs"]) * 1e-4 for h in harmonics_cfg], dtype=np.float64)
        self.phase_arr = np.array([np.deg2rad(float(h["phase_deg"])) for h in harmonics_cfg], dtype=np.float64)
        self.rmin_arr = np.array([float(h["rmin"]) for h in harmonics_cfg], dtype=np.float64)
        self.rmax_arr = np.array([float(h["rmax"]) for h in harmonics_cfg], dtype=np.float64)


def mean_order0_bz(bm, r, theta):
    """Azimuthal mean over ``theta`` of the order-0 Bz = B0 r^k F(theta - s(r))
    at each radius in ``r`` (host numpy; the mean-field curve the closed-orbit
    search starts from)."""
    r = np.atleast_1d(np.asarray(r, dtype=np.float64))
    th = np.asarray(theta, dtype=np.float64)
    period = 2 * np.pi / bm.Nsectors
    s = np.polyval(bm.s_rad_polynomial, r)
    fs = np.mod(th[None, :] - s[:, None], period)
    F = np.zeros_like(fs)
    for x0, x1, w0, mag, sgn in np.asarray(bm.convex_params, dtype=np.float64).reshape(-1, 5):
        ks = 16.0 / w0
        F += sgn * abs(mag) / (1.0 + np.exp(-ks * (fs - x0))) / (1.0 + np.exp(ks * (fs - x1)))
    return bm.B0 * r ** bm.k_value * F.mean(axis=1)


def field_at_points(gp, r, fi, z, precision="fp64", device=0):
    """Static magnetic field (Bz, Bfi, Br) at host points, as a (3, n) numpy
    array, evaluated on the GPU the way a tracking tier evaluates it:
    "fp64" is the FP64 path's field (GetBfields_Njit + Bmap_add_all_orders +
    harmonics, src/FFAG_Field.py:872-884, src/FFAG_MathTools.py:1139-1188;
    ffa_bfield_eval mode 0), "fp32" the FP32 tier's (same terms in FP32 on the
    FP32 map; mode 3)."""
    import torch
    if precision not in ("fp64", "fp32"):
        raise ValueError("precision must be 'fp64' or 'fp32'")
    dev = torch.device("cuda", device)
    ctx = _lib.Context(device, torch.cuda.current_stream(dev).cuda_stream)
    f32 = precision == "fp32"
    tay = gp.Bmap.device_taylor(torch.float32 if f32 else torch.float64, dev)
    ctx.check(ctx.lib.ffa_set_machine(ctx.h, build_machine(gp, tay, f32)), "ffa_set_machine")
    cols = [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev) for a in (r, fi, z)]
    n = cols[0].numel()
    if cols[1].numel() != n or cols[2].numel() != n:
        raise ValueError("r, fi and z must have the same length")
    out = torch.empty(3 * max(n, 1), dtype=torch.float64, device=dev)
    ctx.check(ctx.lib.ffa_bfield_eval(ctx.h, _lib._vp(cols[0].data_ptr()), _lib._vp(cols[1].data_ptr()),
                                      _lib._vp(cols[2].data_ptr()), n, _lib._vp(out.data_ptr()), 3 if f32 else 0,
                                      0.0), "ffa_bfield_eval")
    res = out[: 3 * n].view(3, n).cpu().numpy()
    ctx.close()
    return res


def build_machine(gp, taylor_tensor, taylor_f32: bool) -> "_lib.FfaMachine":
    """Pack FFAG_GlobalParameters into the C-ABI machine description."""
    bm = gp.Bmap
    m = _lib.FfaMachine()
    m.taylor = taylor_tensor.data_ptr() if (taylor_tensor is not None and bm.max_order >= 1) else None
    m.taylor_f32 = 1 if taylor_f32 else 0
    m.max_order = int(bm.max_order)
    m.r_size, m.fi_size = int(bm.r_size), int(bm.fi_size)
    m.r_min, m.r_step, m.fi_min, m.fi_step = float(bm.r_min), float(bm.r_step), float(bm.fi_min), float(bm.fi_step)
    m.B0, m.k_value = float(bm.B0), float(bm.k_value)
    m.nsectors = int(bm.Nsectors)
    cp = np.asarray(bm.convex_params, dtype=np.float64).reshape(-1, 5)
    if cp.shape[0] > _lib.FFA_MAX_SEG:
        raise ValueError(f"at most {_lib.FFA_MAX_SEG} field segments are supported")
    m.n_seg = cp.shape[0]
    for j in range(cp.shape[0]):
        for k in range(5):
            m.convex[j][k] = float(cp[j, k])
    sp, spp = np.asarray(bm.s_rad_polynomial, np.float64), np.asarray(bm.s_rad_polynomial_prime, np.float64)
    if sp.size > _lib.FFA_MAX_POLY or spp.size > _lib.FFA_MAX_POLY:
        raise ValueError("spiral polynomial degree too high")
    m.n_spoly, m.n_sppoly = sp.size, spp.size
    for k in range(sp.size):
        m.s_poly[k] = float(sp[k])
    for k in range(spp.size):
        m.sp_poly[k] = float(spp[k])
    bh = gp.BHarmonics
    if bh is not None and bh.enable_flag and bh.Nharm > 0:
        if bh.Nharm > _lib.FFA_MAX_HARM:
            raise ValueError(f"at most {_lib.FFA_MAX_HARM} harmonics are supported")
        m.n_harm = bh.Nharm
        for k in range(bh.Nharm):
            m.harm_m[k] = float(bh.m_arr[k])
            m.harm_amp[k] = float(bh.amp_arr[k])
            m.harm_phase[k] = float(bh.phase_arr[k])
            m.harm_rmin[k] = float(bh.rmin_arr[k])
            m.harm_rmax[k] = float(bh.rmax_arr[k])
    em = gp.Emap
    if em is not None and getattr(em, "GapType", "fixed") == "spiral":
        ng = len(em.gap_azimuths)
        if ng > _lib.FFA_MAX_GAP:
            raise ValueError(f"at most {_lib.FFA_MAX_GAP} RF gaps are supported")
        m.rf_enable = 1 if em.EnableFlag else 0
        m.n_gap = ng
        m.acc_etot = float(em.acc_Etot_uniform)
        for g in range(ng):
            m.gap_theta[g] = float(em.gap_azimuths[g])
        for g in range(min(len(em.rf_shift), _lib.FFA_MAX_GAP)):
            m.rf_shift[g] = float(em.rf_shift[g])
        m.gap_model = 1
        m.emap_nr, m.emap_nphi = int(em.nr), int(em.nphi)
        m.emap_r0, m.emap_dr = float(em.r0), float(em.dr)
        m.emap_phi0, m.emap_dphi = float(em.phi0), float(em.dphi)
        dev = taylor_tensor.device if taylor_tensor is not None else "cuda"
        m.emap = em.device_emap(dev).data_ptr()
    elif em is not None:
        ng = len(em.acc_regions)
        if ng > _lib.FFA_MAX_GAP:
            raise ValueError(f"at most {_lib.FFA_MAX_GAP} RF gaps are supported")
        m.rf_enable = 1 if em.EnableFlag else 0
        m.n_gap = ng
        m.acc_etot = float(em.acc_Etot_uniform)
        for g in range(ng):
            m.gap_theta[g] = float(em.gap_azimuths[g])
            for j in range(4):
                m.gap_px[g][j] = float(em.acc_regions[g][j, 0])
                m.gap_py[g][j] = float(em.acc_regions[g][j, 1])
            for j in range(4):
                m.gap_box[g][j] = float(em.box_bounds[g][j])
        for g in range(min(len(em.rf_shift), _lib.FFA_MAX_GAP)):
            m.rf_shift[g] = float(em.rf_shift[g])
    if gp.Aperture_enable:
        m.aperture_enable = 1
        for k in range(4):
            m.aperture[k] = float(gp.Aperture_m[k])
    m.q, m.c, m.m0, m.E0 = float(gp.q), float(gp.c), float(gp.m0), float(gp.E0)
    return m
