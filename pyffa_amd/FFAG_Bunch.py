"""Particle store of the tracking path, resident on the GPU.

Same constructor, attributes and methods as the reference's ``FFAG_Bunch``
(src/FFAG_Bunch.py:79-409, 1157-1352).  The reference keeps an (N, 16)
float64 row-major ``LocalBunch`` plus ~60 scratch arrays on the host; here
the state lives in HBM as structure-of-arrays torch tensors (one FP64 array
per state column, one flag byte per particle, the SC endpoint fields as
(3, N) planes) and ``LocalBunch`` is a host mirror refreshed on access.

Column layout of the mirror (``FFAG_ManageBunchAttribute``, src/FFAG_Bunch.py:
23-62): 0 r, 1 vr, 2 z, 3 vz, 4 fi, 5 dfi/dt (Boris) or E_total [J] (RK4),
6 inj_t [ns], 7 Inj_flag, 8 Survive, 9 RF_phase, 10-12 Esc (r, z, fi),
13 Bunch_ID, 14 Local_ID, 15 Global_ID.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class FFAG_ManageBunchAttribute:
    """Column names of the particle matrix (src/FFAG_Bunch.py:23-62)."""

    def __init__(self):
        names = ["r", "vr", "z", "vz", "fi", "Ek", "inj_t", "Inj_flag", "Survive", "RF_phase",
                 "Esc_r", "Esc_z", "Esc_fi", "Bunch_ID", "Local_ID", "Global_ID"]
        self.Attribute = {n: i for i, n in enumerate(names)}
        fmt = {"Inj_flag": "%d", "Survive": "%d", "Bunch_ID": "%d", "Local_ID": "%d", "Global_ID": "%d",
               "inj_t": "%.6e", "RF_phase": "%.6e", "Esc_r": "%.6e", "Esc_z": "%.6e", "Esc_fi": "%.6e"}
        self.AttributeFormat = {n: fmt.get(n, "%.8e") for n in names}

    def get_num_attributes(self):
        return len(self.Attribute)

    def get_attribute_names(self):
        return list(self.Attribute.keys())


def fixed_point_bits(n_particles: int) -> int:
    """Fractional bits per axis of the fixed-point CIC weights: at most 20,
    and small enough that n * 2^(2 bits) < 2^63 so the int64 grid cannot
    overflow even if every particle lands in one cell."""
    b = 20
    while b > 1 and n_particles * (1 << (2 * b)) >= (1 << 63):
        b -= 1
    return b


class FFAG_Bunch:
    """Device-resident bunch (src/FFAG_Bunch.py:79-262).

    Extra keyword arguments (not in the reference): ``device`` (CUDA/HIP
    ordinal) and ``precision`` ('fp64' for the FP64 path, 'fp32' for the
    FP32-field tier, whose SC endpoint fields are stored in float32)."""

    def __init__(self, ParticlesDistribution, marcosize=1, sc_grid_size=(64, 128, 128, 12.0), BunchType="Boris",
                 *, device=0, precision="fp64"):
        import torch
        if BunchType not in ("Boris", "RK4"):
            raise ValueError("BunchType must be Boris or RK4")
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        if not torch.cuda.is_available():
            raise _lib.FfaError("FFAG_Bunch needs a GPU: the tracking path runs only on the HIP kernels")
        self.torch = torch
        self.BunchAttribute = FFAG_ManageBunchAttribute()
        self.BunchType = BunchType
        self.precision = precision
        P = np.ascontiguousarray(ParticlesDistribution, dtype=np.float64)
        if P.ndim != 2 or P.shape[1] < 16:
            raise ValueError("ParticlesDistribution must be an (N, 16) array")
        # a restart file row carries the dump time as a 17th column
        P = np.ascontiguousarray(P[:, :16])
        self.TotalParticleNum = P.shape[0]
        self.TotalParticleNum_Local = P.shape[0]
        self.marcosize = marcosize
        self.speed_c = 2.99792458e8
        self.q_charge = 1.60217662e-19
        self.E0_rest_MeV = 938.2723
        self.StaticBunchLocal = P.copy()
        self.StaticBunchLocal[:, 14] = np.arange(P.shape[0])
        self.LocalUnInjectBunch = self.StaticBunchLocal
        self._uninject_head = 0
        self.uninject_num = self.TotalParticleNum_Local
        self.InjectedNum_Local = 0
        self.SurvivedNum_Local = 0
        self.LostNum_Local = 0
        self.InjectedNum = 0
        self.SurvivedNum = 0
        self.LostNum = 0
        self.f_grid_size = int(sc_grid_size[0])
        self.x_grid_size = int(sc_grid_size[1])
        self.z_grid_size = int(sc_grid_size[2])
        self.sigma_ratio = float(sc_grid_size[3])
        self.SEO_fi_axis = None
        self.SEO_Ek_axis = None
        self.SEO_r_matrix = None
        self.SEO_pr_matrix = None
        self.SEO_perimeter = None
        self.device = torch.device("cuda", device)
        self.last_sc_s = 0.0
        self._upload(self.StaticBunchLocal)
        self._seo_dev = None
        self._sc_ws = None
        self._sc_layout = None
        self._pending_gather = None     # SC params of a solve whose gather the next push applies
        self._host_cache = None
        # store per-cell CIC base counts at each solve (a checking aid)
        self.sc_cell_counts = False
        self.sc_charge_grid = False     # store the int64 fixed-point charge grid of each solve
        self.sc_field_cubes = False     # store the (Ex, Ez) cubes of each solve

    # ------------------------------------------------------------------ state
    def _upload(self, M):
        t = self.torch
        dev = self.device
        n = M.shape[0]
        cols = lambda j: t.from_numpy(np.ascontiguousarray(M[:, j])).to(dev)
        self.r, self.vr, self.z, self.vz = cols(0), cols(1), cols(2), cols(3)
        self.fi, self.fidot, self.rf_phase = cols(4), cols(5), cols(9)
        fl = np.zeros(n, np.uint8)
        fl |= np.where(M[:, 7] == 1, _lib.F_INJ, 0).astype(np.uint8)
        fl |= np.where(M[:, 8] != 0, _lib.F_ALIVE, 0).astype(np.uint8)
        self.flags = t.from_numpy(fl).to(dev)
        edt = t.float32 if self.precision == "fp32" else t.float64
        esc = t.from_numpy(np.ascontiguousarray(M[:, 10:13].T)).to(dev, edt).contiguous()
        self.esc_pre = esc
        self.esc_prepre = esc.clone()
        # input index of the particle held at each position (the FP32 tier
        # reorders particles by azimuth; the host mirror restores the order)
        self.pid = t.arange(n, dtype=t.int64, device=dev)
        self._alt = None
        self._sort_ws = None
        self._meta = M[:, [6, 13, 14, 15]].copy()
        self._host_cache = None
        self._pre_fi = self._pre_flags = None
        self._pre_valid = False
        self._det_idx = self._det_count = None

    _STATE = ("r", "vr", "z", "vz", "fi", "fidot", "rf_phase", "flags", "esc_pre", "esc_prepre", "pid")

    def _particles_of(self, a) -> "_lib.FfaParticles":
        p = _lib.FfaParticles()
        p.r, p.vr, p.z, p.vz = a["r"].data_ptr(), a["vr"].data_ptr(), a["z"].data_ptr(), a["vz"].data_ptr()
        p.fi, p.fidot, p.rf_phase = a["fi"].data_ptr(), a["fidot"].data_ptr(), a["rf_phase"].data_ptr()
        p.flags = a["flags"].data_ptr()
        p.esc_pre, p.esc_prepre = a["esc_pre"].data_ptr(), a["esc_prepre"].data_ptr()
        p.n = self.TotalParticleNum_Local
        p.esc_f32 = 1 if self.precision == "fp32" else 0
        return p

    def particles(self) -> "_lib.FfaParticles":
        return self._particles_of({k: getattr(self, k) for k in self._STATE})

    def sort_by_azimuth(self, ctx=None, nbins=65535):
        """Reorder the injected particles by azimuth bin (stable; see
        ffa_sort_particles).  The host mirror and ``ordered()`` keep
        presenting particles in the input order."""
        if self._pending_gather is not None:
            raise RuntimeError("an SC solve's gather is pending: push before reordering the particles")
        count = self._uninject_head
        if count < 2:
            return
        if ctx is None:
            ctx = self._own_ctx()
        t = self.torch
        if self._alt is None:
            self._alt = {k: t.empty_like(getattr(self, k)) for k in self._STATE}
        if self._sort_ws is None or self._sort_ws[0] < count:
            nbytes = ctypes.c_size_t()
            ctx.check(ctx.lib.ffa_sort_workspace_size(count, ctypes.byref(nbytes)), "ffa_sort_workspace_size")
            self._sort_ws = (count, t.empty(max(nbytes.value, 1), dtype=t.uint8, device=self.device))
        alt = self._alt
        ctx.check(ctx.lib.ffa_sort_particles(ctx.h, self.particles(), self._particles_of(alt), count, int(nbins),
                                             _lib._vp(self.pid.data_ptr()), _lib._vp(alt["pid"].data_ptr()),
                                             _lib._vp(self._sort_ws[1].data_ptr())), "ffa_sort_particles")
        cur = {k: getattr(self, k) for k in self._STATE}
        for k in self._STATE:
            setattr(self, k, alt[k])
        self._alt = cur
        self._host_cache = None
        self._pre_valid = False         # recorded pre-step arrays follow the old order

    def ordered(self, name):
        """Device copy of a state array in the input particle order."""
        a = getattr(self, name)
        out = self.torch.empty_like(a)
        if a.dim() == 2:
            out[:, self.pid] = a
        else:
            out[self.pid] = a
        return out

    def invalidate_host(self):
        self._host_cache = None

    def _assemble(self, pos=None):
        """(k, 16) host matrix of the particles held at device positions
        ``pos`` (an int64 device tensor), or of all particles in the input
        order when ``pos`` is None."""
        if self._pending_gather is not None:
            raise RuntimeError("an SC solve's gather is pending: the endpoint fields are not written yet")
        t = self.torch
        if pos is None:
            take = self.ordered
            meta = self._meta
        else:
            def take(name):
                a = getattr(self, name)
                return a[:, pos] if a.dim() == 2 else a[pos]
            meta = self._meta[self.pid[pos].cpu().numpy()]
        M = np.zeros((meta.shape[0], 16))
        for j, name in ((0, "r"), (1, "vr"), (2, "z"), (3, "vz"), (4, "fi"), (5, "fidot"), (9, "rf_phase")):
            M[:, j] = take(name).cpu().numpy()
        fl = take("flags").cpu().numpy()
        M[:, 7] = (fl & _lib.F_INJ) != 0
        M[:, 8] = (fl & _lib.F_ALIVE) != 0
        M[:, [6, 13, 14, 15]] = meta
        pre = take("esc_pre").to(t.float64).cpu().numpy()
        pp = take("esc_prepre").to(t.float64).cpu().numpy()
        both = ((fl & _lib.F_SCPRE) != 0) & ((fl & _lib.F_SCPREPRE) != 0)
        esc = pre.copy()
        esc[:, both] = pre[:, both] + self.last_sc_s * (pre[:, both] - pp[:, both])
        M[:, 10:13] = esc.T
        return M

    @property
    def LocalBunch(self):
        """(N, 16) host mirror of the device state (read-only view of a copy;
        call ``set_local_bunch`` to write a modified matrix back)."""
        if self._host_cache is None:
            self._host_cache = self._assemble()
        return self._host_cache

    def rows(self, input_rows):
        """Rows ``input_rows`` (indices in the input order) of the host
        mirror, gathered on the device so only those rows are copied."""
        t = self.torch
        idx = t.as_tensor(np.asarray(input_rows, dtype=np.int64), device=self.device)
        return self._assemble(self._positions(idx))

    def _positions(self, input_rows):
        """Device positions of the particles with the given input indices."""
        t = self.torch
        inv = t.empty_like(self.pid)
        inv[self.pid] = t.arange(self.pid.numel(), dtype=t.int64, device=self.device)
        return inv[input_rows]

    def global_ids(self):
        """Global_ID column (15) of the particle matrix, input order."""
        return self._meta[:, 3].astype(np.int64)

    def set_local_bunch(self, M):
        """Replace the device state by an (N, 16) matrix."""
        M = np.ascontiguousarray(M, dtype=np.float64)
        if M.shape != (self.TotalParticleNum_Local, 16):
            raise ValueError("matrix shape mismatch")
        self._upload(M)

    @property
    def Step_SurviveFlag(self):
        """Survive flags (column 8) as they were before the step, when
        ``UpdatePreSteps`` recorded them; otherwise the current flags."""
        fl = self._pre_flags if self._pre_valid else self.flags
        out = self.torch.empty_like(fl)
        out[self.pid] = fl
        return (out.cpu().numpy() & _lib.F_ALIVE) != 0

    @property
    def PreStep_Fi(self):
        if not self._pre_valid:
            raise RuntimeError("UpdatePreSteps() has not recorded a pre-step state")
        out = self.torch.empty_like(self._pre_fi)
        out[self.pid] = self._pre_fi
        return out.cpu().numpy()

    @property
    def PostStep_Fi(self):
        return self.ordered("fi").cpu().numpy()

    # ----------------------------------------------------------- bookkeeping
    def injection_range(self, t_threshold, t_step):
        """Index range injected at this step: inj_t [ns] in [t - dt, t]
        searched over the not-yet-injected tail (src/FFAG_Bunch.py:265-303)."""
        hi = t_threshold * 1e9
        lo = (t_threshold - t_step) * 1e9
        tail = self.LocalUnInjectBunch[self._uninject_head:, 6]
        a = int(np.searchsorted(tail, lo, side="left"))
        b = int(np.searchsorted(tail, hi, side="right"))
        if b - a <= 0:
            return 0, 0
        return self._uninject_head + a, self._uninject_head + b

    def _account_injection(self, begin, end):
        k = end - begin
        if k > 0:
            self._uninject_head = end
            self.uninject_num -= k
            self.InjectedNum_Local += k
            self.SurvivedNum_Local += k

    def InjectParticles(self, t_threshold, t_step, ctx=None):
        b, e = self.injection_range(t_threshold, t_step)
        if e > b:
            if ctx is None:
                ctx = self._own_ctx()
            ctx.check(ctx.lib.ffa_inject(ctx.h, self.particles(), b, e), "ffa_inject")
            self._account_injection(b, e)
            self._host_cache = None
        return 0

    def _own_ctx(self):
        if getattr(self, "_ctx", None) is None:
            self._ctx = _lib.Context(self.device.index or 0,
                                     self.torch.cuda.current_stream(self.device).cuda_stream)
        return self._ctx

    def restore_from_restart_matrix(self, *, assert_monotonic=True):
        """Rebuild the injection head and counters from the flags
        (src/FFAG_Bunch.py:305-366)."""
        M = self.LocalBunch
        n = M.shape[0]
        inj = M[:, 7] > 0.5
        sur = M[:, 8] > 0.5
        zeros = np.flatnonzero(~inj)
        head = int(zeros[0]) if zeros.size else n
        if assert_monotonic and not (inj[:head].all() and (~inj[head:]).all()):
            raise RuntimeError("Restart bunch violates the designed structure: Inj_flag is not a step function "
                               "along inj_t-sorted LocalBunch.")
        self._uninject_head = head
        self.InjectedNum_Local = int(inj.sum())
        self.uninject_num = n - self.InjectedNum_Local
        self.SurvivedNum_Local = int(sur.sum())
        self.LostNum_Local = int((inj & ~sur).sum())

    def UpdateParticleNumGlobal(self):
        self.InjectedNum = self.InjectedNum_Local
        self.SurvivedNum = self.SurvivedNum_Local
        self.LostNum = self.LostNum_Local

    def UpdatePreSteps(self):
        """Record the azimuth and flags before a step on the device
        (PreStep_Fi and Step_SurviveFlag, src/FFAG_Bunch.py:395-401 and
        src/FFAG_MathTools.py:2179-2201).  The push kernels read the flags
        directly, so the tracker calls this only before the steps that a
        detector examines."""
        t = self.torch
        n = self.TotalParticleNum_Local
        if self._pre_fi is None or self._pre_fi.numel() != n:
            self._pre_fi = t.empty(n, dtype=t.float64, device=self.device)
            self._pre_flags = t.empty(n, dtype=t.uint8, device=self.device)
        self._pre_fi.copy_(self.fi)
        self._pre_flags.copy_(self.flags)
        self._pre_valid = True
        self._host_cache = None

    def UpdatePostSteps(self):
        self._host_cache = None

    def detect_crossing(self, target, input_rows=None, ctx=None):
        """detect_azimuth_crossing (src/FFAG_MathTools.py:1633-1695) between
        the state recorded by ``UpdatePreSteps`` and the current one, on the
        GPU (ffa_detect_crossing).

        ``target`` is the azimuth in [0, 2 pi] after the reference's
        ``% (2 pi)``.  Without ``input_rows``: returns the crossing particles'
        input indices in increasing order and their device positions.  With
        ``input_rows`` (the detect_azimuth_crossing_portion form,
        src/FFAG_MathTools.py:1697-1757): returns a boolean mask over them."""
        t = self.torch
        if not self._pre_valid:
            raise RuntimeError("UpdatePreSteps() must run before the step a detector examines")
        if ctx is None:
            ctx = self._own_ctx()
        if input_rows is None:
            pos = None
            fi_pre, fl_pre, fi_post = self._pre_fi, self._pre_flags, self.fi
        else:
            rows = t.as_tensor(np.asarray(input_rows, dtype=np.int64), device=self.device)
            pos = self._positions(rows)
            fi_pre, fl_pre, fi_post = self._pre_fi[pos], self._pre_flags[pos], self.fi[pos]
        n = fi_pre.numel()
        if not target < 2.0 * np.pi:
            # float rounding of x % (2 pi) can give 2 pi itself: no azimuth
            # reduced to [0, 2 pi) reaches it, so nothing crosses
            hit = t.empty(0, dtype=t.int64, device=self.device)
        else:
            if self._det_idx is None or self._det_idx.numel() < max(n, 1):
                self._det_idx = t.empty(max(n, 1), dtype=t.int64, device=self.device)
                self._det_count = t.zeros(1, dtype=t.int64, device=self.device)
            ctx.check(ctx.lib.ffa_detect_crossing(
                ctx.h, _lib._vp(fi_pre.data_ptr()), _lib._vp(fl_pre.data_ptr()), _lib._vp(fi_post.data_ptr()),
                n, float(target), _lib._vp(self._det_idx.data_ptr()), _lib._vp(self._det_count.data_ptr())),
                "ffa_detect_crossing")
            hit = self._det_idx[:int(self._det_count.item())]
        if input_rows is not None:
            mask = np.zeros(n, dtype=bool)
            mask[hit.cpu().numpy()] = True
            return mask
        rows = self.pid[hit]
        order = t.argsort(rows)
        return rows[order].cpu().numpy(), hit[order]

    # ---------------------------------------------------------------- SC data
    def seo_device(self):
        if self._seo_dev is None:
            t = self.torch
            if self.SEO_r_matrix is None:
                raise ValueError("SEO tables (SEO_fi_axis, SEO_Ek_axis, SEO_r_matrix, SEO_pr_matrix, "
                                 "SEO_perimeter) must be set before a space-charge solve")
            f = lambda a: t.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(self.device)
            self._seo_dev = dict(r=f(self.SEO_r_matrix), pr=f(self.SEO_pr_matrix), ek=f(self.SEO_Ek_axis),
                                 per=f(self.SEO_perimeter))
        return self._seo_dev

    def sc_workspace(self):
        if self._sc_ws is None:
            L = _lib.sc_layout(self.TotalParticleNum_Local, self.f_grid_size, self.z_grid_size, self.x_grid_size)
            self._sc_layout = L
            self._sc_ws = self.torch.zeros(L.total, dtype=self.torch.uint8, device=self.device)
        return self._sc_ws, self._sc_layout

    def sc_params(self, *, first, s_last):
        seo = self.seo_device()
        sp = _lib.FfaScParams()
        sp.ns, sp.nx, sp.nz = self.f_grid_size, self.x_grid_size, self.z_grid_size
        sp.sigma_ratio = self.sigma_ratio
        sp.q_macro = self.marcosize * self.q_charge
        sp.rk = 1 if self.BunchType == "RK4" else 0
        sp.first = 1 if first else 0
        sp.s_last = float(s_last)
        sp.bits = fixed_point_bits(max(self.TotalParticleNum_Local, 1))
        sp.cell_counts = 1 if self.sc_cell_counts else 0
        sp.charge_grid = 1 if self.sc_charge_grid else 0
        sp.field_cubes = 1 if self.sc_field_cubes else 0
        sp.seo_r, sp.seo_pr = seo["r"].data_ptr(), seo["pr"].data_ptr()
        sp.seo_ek_axis, sp.seo_perimeter = seo["ek"].data_ptr(), seo["per"].data_ptr()
        sp.n_ek, sp.n_fi = int(self.SEO_r_matrix.shape[0]), int(self.SEO_r_matrix.shape[1])
        sp.ek0 = float(self.SEO_Ek_axis[0])
        sp.dek = float(self.SEO_Ek_axis[1] - self.SEO_Ek_axis[0])
        sp.fi0 = float(self.SEO_fi_axis[0])
        sp.dfi = float(self.SEO_fi_axis[1] - self.SEO_fi_axis[0])
        sp.E0_MeV = self.E0_rest_MeV
        sp.c = self.speed_c
        return sp

    def sc_view(self, name):
        """Typed device view of one workspace region (after a solve); "ex"
        and "ez" are the components of the interleaved field region, written
        by solves with ``sc_field_cubes`` set (and by every solve on grids
        above 2^23 cells); "phi" is written by every solve except those on
        grids above 2^23 cells without ``sc_field_cubes``; "grid" is written
        only by solves with ``sc_charge_grid`` set."""
        t = self.torch
        if name in ("ex", "ez"):
            return self.sc_view("exz")[..., 0 if name == "ex" else 1]
        ws, L = self.sc_workspace()
        ns, nz, nx, n = self.f_grid_size, self.z_grid_size, self.x_grid_size, self.TotalParticleNum_Local
        spec = {"grid": (t.int64, (ns, nz, nx)), "cell_count": (t.int32, (ns, nz, nx)),
                "phi": (t.float32, (ns, nz, nx)), "exz": (t.float32, (ns, nz, nx, 2)),
                "px": (t.float64, (n,)), "ppr0": (t.float64, (n,)), "pslice": (t.int32, (n,)),
                "slice_count": (t.int32, (ns,)), "result": (t.float64, (32,))}[name]
        off = getattr(L, name)
        dt, shape = spec
        cnt = int(np.prod(shape))
        nbytes = cnt * t.empty((), dtype=dt).element_size()
        return ws[off: off + nbytes].view(dt).view(shape)

    def sc_result(self):
        r = self.sc_view("result").cpu().numpy()
        names = [f[0] for f in _lib.FfaScResult._fields_]
        out = {}
        k = 0
        for nm, ty in _lib.FfaScResult._fields_:
            size = ctypes_len(ty)
            out[nm] = r[k] if size == 1 else r[k: k + size].copy()
            k += size
        return out

    def Build_2_5D_SliceGrid_sbs(self, fft_fwd_1=None, fft_inv_1=None, ctx=None, *, first=True, s_last=0.0):
        """One 2.5D SC solve on the device (src/FFAG_Bunch.py:1157-1352).
        The FFT plan arguments are accepted for signature compatibility and
        unused.  Returns the reference's 18-tuple with host (numpy) arrays;
        the tracking loop calls ``sc_solve_device`` instead, which copies
        nothing back."""
        keep = self.sc_field_cubes, self.sc_charge_grid
        self.sc_field_cubes = self.sc_charge_grid = True
        try:
            self.sc_solve_device(ctx, first=first, s_last=s_last)
        finally:
            self.sc_field_cubes, self.sc_charge_grid = keep
        from . import FFAG_SC_new as SC
        res = self.sc_result()
        ns = self.f_grid_size
        bad = res["bad_grid"] != 0.0
        rho = self.sc_view("grid").cpu().numpy().astype(np.float64) * (
            res["qscale"] / float(1 << (2 * fixed_point_bits(max(self.TotalParticleNum_Local, 1)))))
        fi_grid = np.linspace(res["fi_s"], res["fi_e"], ns)
        s_grid = fi_grid / np.pi / 2.0 * res["perimeter"]
        counts = self.sc_view("slice_count").cpu().numpy()
        rho_f = counts * (self.marcosize * self.q_charge)
        if bad:
            # degenerate box: the reference returns zero cubes and a zero
            # longitudinal field, and leaves the particles' fields as they were
            shape = (ns, self.z_grid_size, self.x_grid_size)
            phi, ex, ez = (np.zeros(shape, np.float32) for _ in range(3))
            ef = np.zeros(ns)
        else:
            phi = self.sc_view("phi").cpu().numpy()
            ex = self.sc_view("ex").cpu().numpy()
            ez = self.sc_view("ez").cpu().numpy()
            ef = SC.total_field_1d_numba(s_grid, rho_f)
        # survivors of this solve (its alive flags) in the input order; the
        # solve's x = r - r0 is held in device order, so it is put back in
        # the input order through the particle ids first
        t = self.torch
        alive_dev = (self.flags & _lib.F_ALIVE) != 0
        alive = self.ordered("flags").cpu().numpy() & _lib.F_ALIVE != 0
        px_in = t.zeros_like(self.r)
        px_in[self.pid[alive_dev]] = self.sc_view("px")[alive_dev]
        px = px_in.cpu().numpy()[alive]
        M = self.LocalBunch
        return (rho.astype(np.float32), phi, ex, ez, res["xmin"], res["xmax"], res["zmin"], res["zmax"],
                res["bunch_length"], ef, s_grid, rho_f, M[alive, 0], M[alive, 4], px,
                M[alive, 0] - px, M[alive, 2], fi_grid)

    def sc_solve_device(self, ctx=None, *, first, s_last, defer_gather=False):
        """One SC solve on the device.  defer_gather: stop after the field;
        the gather and endpoint roll then run inside the next FP32 push
        launch (``push_fused_f32``), which must follow before anything reads
        the endpoint fields or flags."""
        if ctx is None:
            ctx = self._own_ctx()
        if self._pending_gather is not None:
            raise RuntimeError("an SC solve's gather is still pending: push before the next solve")
        ws, _ = self.sc_workspace()
        sp = self.sc_params(first=first, s_last=s_last)
        if defer_gather:
            ctx.check(ctx.lib.ffa_sc_solve_field(ctx.h, self.particles(), sp, ws.data_ptr()), "ffa_sc_solve_field")
            self._pending_gather = sp
            return
        ctx.check(ctx.lib.ffa_sc_solve(ctx.h, self.particles(), sp, ws.data_ptr()), "ffa_sc_solve")
        self._finish_sc_roll()

    def _finish_sc_roll(self):
        # the solve writes the new pre-step field into the prepre buffer and
        # leaves the old pre-step field, now the prepre one, where it was
        self.esc_pre, self.esc_prepre = self.esc_prepre, self.esc_pre
        self._host_cache = None

    def push_fused_f32(self, ctx, steps, inj, nsteps):
        """nsteps fused FP32 steps (ffa_push_boris_fused_f32), applying first
        the gather of a solve that deferred it (ffa_push_boris_fused_f32_sc)."""
        sp = self._pending_gather
        if sp is None:
            ctx.check(ctx.lib.ffa_push_boris_fused_f32(ctx.h, self.particles(), steps, inj, nsteps),
                      "ffa_push_boris_fused_f32")
            return
        ws, _ = self.sc_workspace()
        ctx.check(ctx.lib.ffa_push_boris_fused_f32_sc(ctx.h, self.particles(), steps, inj, nsteps, sp,
                                                      ws.data_ptr()), "ffa_push_boris_fused_f32_sc")
        self._pending_gather = None
        self._finish_sc_roll()


def ctypes_len(ty):
    return getattr(ty, "_length_", 1)
