"""Learned field surrogate for the FP32 tier (no counterpart in the reference).

``FFAG_FieldSurrogate`` is a Fourier-feature MLP (32 frequencies -> 64
features, four hidden layers of 128 SiLU units, 3 outputs) that stands in for
the Taylor-map gather and z extrapolation ``Bmap_add_all_orders``
(src/FFAG_MathTools.py:1139-1188) for particles with ``|z| <= zt``.  It is
trained on the GPU, by the library's own forward, backward and Adam kernels,
on samples of the method's own Taylor-extrapolated field: the coefficient
form of the map gather at random points of one sector, evaluated in FP64 by
``ffa_bfield_eval``.  The order-0 analytic field ``GetBfields_Njit`` stays
analytic; outside ``|z| <= zt`` the push falls back to the map gather.

The network's z dependence is built in: it predicts ``o(r, theta, (z/zt)^2)``
and the field is ``dBz = S_z c^2 g(o_0)``, ``dBfi = S_fi c^3 g(o_1)``,
``dBr = S_r c^3 g(o_2)`` with ``c = z / zt``, which keeps the parity of the
Taylor series and the vanishing of the correction on the mid-plane exact.
``g(o) = o`` by default; with ``delta > 0``, ``g(o) = delta sinh(o)`` and the
network is trained on ``asinh(coefficient / (S delta))``, which trades
accuracy at the fringes for accuracy in the magnet bodies.  On the 12-sector
test lattice the linear output tracks the map method best over 2000 steps.

Typical use::

    sur = FFAG_FieldSurrogate(gp_order4, zt=0.01)   # gp_order4.Bmap has max_order 4
    sur.build_dataset(4_000_000)
    sur.train(iters=12000)
    print(sur.field_error())
    LocalParameters["surrogate"] = sur               # FP32 tier uses it
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from .FFAG_Field import build_machine

# azimuthal harmonics of the sector (geometric ladder 1.3^j, duplicates removed)
_HARMONICS = (1, 2, 3, 4, 5, 6, 8, 11, 14, 18, 23, 30, 39, 51, 67, 87, 112, 146, 190, 247, 321, 418)
# radial / z^2 frequencies of the remaining features, in quarter revolutions
_RZ = ((0.5, 0.0), (1.0, 0.0), (2.0, 0.0), (0.5, 1.0), (1.0, 1.0), (2.0, 1.0), (1.0, 2.0), (3.0, 0.0),
       (0.5, 0.0), (1.0, 0.0))


def default_frequencies():
    """(m, wa, wc) of the 32 Fourier features, wa and wc in revolutions."""
    m = np.zeros(_lib.MLP_NFREQ, np.float32)
    wa = np.zeros(_lib.MLP_NFREQ, np.float32)
    wc = np.zeros(_lib.MLP_NFREQ, np.float32)
    m[: len(_HARMONICS)] = _HARMONICS
    for q, (a, c) in enumerate(_RZ):
        wa[len(_HARMONICS) + q] = a / 4.0
        wc[len(_HARMONICS) + q] = c / 4.0
    return m, wa, wc


def init_params(seed=0):
    """Uniform He-style initialisation of the FP32 master vector (W stored
    [in][out], order W0 b0 ... W4 b4)."""
    rng = np.random.default_rng(seed)
    sizes = [(64, 128), (128, 128), (128, 128), (128, 128), (128, 3)]
    out = []
    for fi, fo in sizes:
        lim = math.sqrt(6.0 / fi) * (0.5 if fo == 3 else 1.0)
        out.append(rng.uniform(-lim, lim, fi * fo).astype(np.float32))
        out.append(np.zeros(fo, np.float32))
    p = np.concatenate(out)
    assert p.size == _lib.MLP_NPARAM
    return p


class FFAG_FieldSurrogate:
    def __init__(self, GlobalParameters, zt=0.01, seed=0, device="cuda", fallback=True, ctx=None, delta=0.0):
        import torch
        self.torch = torch
        self.gp = GlobalParameters
        bm = GlobalParameters.Bmap
        if bm.max_order < 1:
            raise ValueError("the surrogate learns the Taylor-map part: the Bmap needs max_order >= 1")
        self.device = torch.device(device)
        self.zt = float(zt)
        self.r_min = float(bm.r_min)
        self.r_max = float(bm.r_min + bm.r_step * (bm.r_size - 1))
        self.sector = 2.0 * np.pi / float(bm.Nsectors)
        self.freq = default_frequencies()
        self.scale = np.ones(3)
        self.delta = float(delta)
        self.fallback = bool(fallback)
        self.ctx = ctx or _lib.Context(self.device.index or 0, torch.cuda.current_stream(self.device).cuda_stream)
        self._machine_tay = None
        p = init_params(seed)
        self.params = torch.tensor(p, device=self.device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.packed = torch.zeros(_lib.MLP_PACKED_BYTES, dtype=torch.uint8, device=self.device)
        self.step = 0
        self._ws = None
        self._ws_n = 0
        self.loss_history = []
        self._pack()

    # ------------------------------------------------------------- plumbing
    def desc(self) -> _lib.FfaMlpDesc:
        d = _lib.FfaMlpDesc()
        d.r_c = 0.5 * (self.r_min + self.r_max)
        d.inv_rh = 2.0 / (self.r_max - self.r_min)
        d.inv_sector = 1.0 / self.sector
        d.zt = self.zt
        for k in range(3):
            d.out_scale[k] = float(self.scale[k])
        d.out_delta = self.delta
        m, wa, wc = self.freq
        for k in range(_lib.MLP_NFREQ):
            d.m[k], d.wa[k], d.wc[k] = float(m[k]), float(wa[k]), float(wc[k])
        d.fallback = 1 if self.fallback else 0
        return d

    def _check(self, rc, what):
        self.ctx.check(rc, what)

    def _pack(self):
        self._check(self.ctx.lib.ffa_mlp_pack(self.ctx.h, _lib._vp(self.params.data_ptr()),
                                              _lib._vp(self.packed.data_ptr())), "ffa_mlp_pack")

    def _workspace(self, n):
        if self._ws is None or self._ws_n != n:
            L = _lib.mlp_layout(n)
            self._ws = self.torch.empty(L.total, dtype=self.torch.uint8, device=self.device)
            self._ws_n = n
        return self._ws

    def _set_label_machine(self):
        """FP64 map of the Bmap in the context (labels and the judge)."""
        tay = self.gp.Bmap.device_taylor(self.torch.float64, self.device)
        self._machine_tay = tay
        m = build_machine(self.gp, tay, False)
        self._check(self.ctx.lib.ffa_set_machine(self.ctx.h, m), "ffa_set_machine")

    def bfield(self, r, fi, z, mode):
        """ffa_bfield_eval on FP64 device tensors -> (3, n) FP64."""
        n = r.numel()
        out = self.torch.empty(3, n, dtype=self.torch.float64, device=self.device)
        self._check(self.ctx.lib.ffa_bfield_eval(self.ctx.h, _lib._vp(r.data_ptr()), _lib._vp(fi.data_ptr()),
                                                 _lib._vp(z.data_ptr()), n, _lib._vp(out.data_ptr()), mode,
                                                 self.zt), "ffa_bfield_eval")
        return out

    def sample_points(self, n, seed, z_range=None):
        """Uniform points over r_min..r_max, one sector, |z| <= z_range (FP64)."""
        torch = self.torch
        g = torch.Generator(device=self.device)
        g.manual_seed(int(seed))
        zr = self.zt if z_range is None else float(z_range)
        r = self.r_min + (self.r_max - self.r_min) * torch.rand(n, dtype=torch.float64, device=self.device, generator=g)
        fi = self.sector * torch.rand(n, dtype=torch.float64, device=self.device, generator=g)
        z = zr * (2.0 * torch.rand(n, dtype=torch.float64, device=self.device, generator=g) - 1.0)
        return r, fi, z

    # ---------------------------------------------------------------- data
    def build_dataset(self, n=4_000_000, seed=1):
        """Samples labelled with the map's Taylor part in coefficient form;
        sets the output scales to the label rms."""
        torch = self.torch
        self._set_label_machine()
        r, fi, z = self.sample_points(n, seed)
        coef = self.bfield(r, fi, z, 2)
        self.scale = torch.sqrt((coef * coef).mean(dim=1)).cpu().numpy()
        self.scale[self.scale == 0] = 1.0
        s = torch.tensor(self.scale, dtype=torch.float64, device=self.device)[:, None]
        self.pts = torch.stack([r, fi, z]).float().contiguous()
        t = coef / s
        if self.delta > 0:
            t = torch.asinh(t / self.delta)
        self.tgt = t.float().contiguous()
        self.n_data = n
        return self

    # --------------------------------------------------------------- train
    def grad(self, pts, tgt):
        """(grad, loss) of the MSE over a batch (3, n) float32 device tensors."""
        torch = self.torch
        n = pts.shape[1]
        g = torch.empty_like(self.params)
        loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        ws = self._workspace(n)
        self._check(self.ctx.lib.ffa_mlp_grad(self.ctx.h, ctypes.byref(self.desc()), _lib._vp(self.packed.data_ptr()),
                                              _lib._vp(pts.data_ptr()), _lib._vp(tgt.data_ptr()), n,
                                              _lib._vp(ws.data_ptr()), _lib._vp(g.data_ptr()),
                                              _lib._vp(loss.data_ptr())), "ffa_mlp_grad")
        return g, loss

    def adam(self, grad, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        hp = _lib.FfaAdam(lr=lr, beta1=beta1, beta2=beta2, eps=eps, step=self.step)
        self._check(self.ctx.lib.ffa_mlp_adam(self.ctx.h, _lib._vp(self.params.data_ptr()), _lib._vp(grad.data_ptr()),
                                              _lib._vp(self.m.data_ptr()), _lib._vp(self.v.data_ptr()),
                                              _lib._vp(self.packed.data_ptr()), ctypes.byref(hp)), "ffa_mlp_adam")

    def train(self, iters=12000, batch=8192, lr=2e-3, lr_final=0.0, seed=2, log_every=500, verbose=False):
        """Adam with a cosine learning-rate schedule over ``iters`` steps on
        reshuffled passes over the dataset; returns the loss history
        [(iteration, loss)]."""
        torch = self.torch
        if not hasattr(self, "pts"):
            raise RuntimeError("build_dataset() first")
        n = self.n_data
        nb = n // batch
        if nb < 1:
            raise ValueError("batch larger than the dataset")
        g = torch.Generator(device=self.device)
        g.manual_seed(int(seed))
        ws = self._workspace(batch)
        desc = self.desc()
        loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        hist = []
        blocks_p = blocks_t = None
        for it in range(iters):
            k = it % nb
            if k == 0:
                perm = torch.randperm(n, device=self.device, generator=g)[: nb * batch]
                blocks_p = self.pts[:, perm].reshape(3, nb, batch).permute(1, 0, 2).contiguous()
                blocks_t = self.tgt[:, perm].reshape(3, nb, batch).permute(1, 0, 2).contiguous()
            self.step += 1
            lr_t = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * it / max(1, iters)))
            hp = _lib.FfaAdam(lr=lr_t, beta1=0.9, beta2=0.999, eps=1e-8, step=self.step)
            self._check(self.ctx.lib.ffa_mlp_train_step(
                self.ctx.h, ctypes.byref(desc), _lib._vp(self.params.data_ptr()), _lib._vp(self.m.data_ptr()),
                _lib._vp(self.v.data_ptr()), _lib._vp(self.packed.data_ptr()), _lib._vp(blocks_p[k].data_ptr()),
                _lib._vp(blocks_t[k].data_ptr()), batch, _lib._vp(ws.data_ptr()), ctypes.byref(hp),
                _lib._vp(loss.data_ptr())), "ffa_mlp_train_step")
            if log_every and (it % log_every == 0 or it == iters - 1):
                hist.append((it, float(loss.item())))
                if verbose:
                    print(f"[surrogate] iter {it} loss {hist[-1][1]:.3e}", flush=True)
        self.loss_history += hist
        return hist

    # ------------------------------------------------------------ evaluate
    def eval(self, pts, mode=0):
        """Network field (mode 0: dBz, dBfi, dBr) or raw outputs (mode 1) at
        (3, n) float32 points (r, fi, z)."""
        n = pts.shape[1]
        out = self.torch.empty(3, n, dtype=self.torch.float32, device=self.device)
        self._check(self.ctx.lib.ffa_mlp_eval(self.ctx.h, ctypes.byref(self.desc()), _lib._vp(self.packed.data_ptr()),
                                              _lib._vp(pts.data_ptr()), n, _lib._vp(out.data_ptr()), mode),
                    "ffa_mlp_eval")
        return out

    def field_error(self, n=1_000_000, seed=7, z_range=None):
        """Error of (order-0 field + surrogate) against the FP64 map method,
        relative to the rms total field, on fresh uniform points; beside it
        the error of the order-0 field alone (the method without maps)."""
        torch = self.torch
        self._set_label_machine()
        r, fi, z = self.sample_points(n, seed, z_range)
        total = self.bfield(r, fi, z, 0)
        taylor = self.bfield(r, fi, z, 1)
        pred = self.eval(torch.stack([r, fi, z]).float().contiguous(), 0).double()
        inside = (z.abs() <= self.zt)
        pred = torch.where(inside[None, :], pred, taylor)     # push fallback outside the trained range
        den = torch.sqrt((total * total).sum(0).mean())
        err = torch.sqrt(((pred - taylor) ** 2).sum(0)) / den
        err0 = torch.sqrt((taylor * taylor).sum(0)) / den
        q = torch.tensor([0.99], dtype=torch.float64, device=self.device)
        w = int(torch.argmax(err))
        worst = dict(r=float(r[w]), fi_in_sector_deg=float(np.rad2deg(float(fi[w]) % self.sector)), z=float(z[w]),
                     B_total=[float(v) for v in total[:, w]], taylor_part=[float(v) for v in taylor[:, w]],
                     network=[float(v) for v in pred[:, w]])
        return dict(points=n, z_range=float(z_range or self.zt), rms_B=float(den),
                    rel_rms=float(torch.sqrt((err * err).mean())), rel_p99=float(torch.quantile(err[:1_000_000], q)),
                    rel_max=float(err.max()), worst_point=worst,
                    order0_rel_rms=float(torch.sqrt((err0 * err0).mean())),
                    order0_rel_p99=float(torch.quantile(err0[:1_000_000], q)), order0_rel_max=float(err0.max()))

    # ------------------------------------------------------------- push use
    def enable(self, ctx):
        """Make ctx's FP32 push use this network (weights must stay alive)."""
        ctx.check(ctx.lib.ffa_mlp_set(ctx.h, ctypes.byref(self.desc()), _lib._vp(self.packed.data_ptr())),
                  "ffa_mlp_set")

    @staticmethod
    def disable(ctx):
        ctx.check(ctx.lib.ffa_mlp_set(ctx.h, None, None), "ffa_mlp_set")

    def state_dict(self):
        return dict(params=self.params.cpu().numpy(), scale=np.asarray(self.scale), zt=self.zt, delta=self.delta,
                    freq=[np.asarray(f) for f in self.freq], step=self.step)

    def load_state_dict(self, sd):
        self.params.copy_(self.torch.tensor(np.asarray(sd["params"], np.float32), device=self.device))
        self.scale = np.asarray(sd["scale"], np.float64)
        self.zt = float(sd["zt"])
        self.delta = float(sd.get("delta", 0.0))
        self.freq = tuple(np.asarray(f, np.float32) for f in sd["freq"])
        self.step = int(sd.get("step", 0))
        self._pack()
