"""Taylor-coefficient maps of the mid-plane field, generated on the GPU.

Same entry points as the reference's map generator:

* ``Generate_Bmaps_MultiProcess(MaxOrder, r_axis, fi_axis_rad, convex_paras,
  s_values_FitCoef_rad, BzR_coef_k)`` -> ``(Br, Bz, Bfi)`` lists of
  ``(len(r_axis), len(fi_axis_rad))`` arrays, orders 0..MaxOrder
  (src/FFAG_HighOrder.py:936-998);
* ``main_Bmap(json_path)`` -- the examples/main_Bmap.py flow: reads
  ``config_Bmap.json``, writes ``Bz/Br/Bfi_order_<n>.npz`` and the
  ``config_Bmap.json`` (with ``Bmean`` and ``r_axis``) the tracker loads
  (examples/main_Bmap.py:17-81, src/FFAG_Field.py:904-943);
* ``calculate_s`` and ``BmapParams_to_ConvexParams`` (:400-421, :1001-1036).

For B(r, theta) = A(r) G(theta - s(r)), with A = a r^k (or a power series)
and G the sum of sigmoid-edged poles, the maps are
  Bz_n = (-1)^n/(2n)! Lap^n B,   Br_n = (-1)^n/(2n+1)! d/dr Lap^n B,
  Bfi_n = (-1)^n/(2n+1)! (1/r) d/dtheta Lap^n B.
Expanding the polar Laplacian (sum of c r^-p d^a_r d^b_theta), the Leibniz
rule on A G and the chain rule on G(theta - s(r)) turns each map row into
  map(r_i, theta_j) = sum_q w_q(r_i) G^(q)(theta_j - s(r_i)),
with scalar weights w_q(r) built from A, s and their derivatives.  The host
computes those weights for all radii at once (numpy over r); the GPU
tabulates G^(q) on the azimuth grid and fills every map value
(``ffa_taylor_maps``).  As in the reference, G^(q)(theta - s(r)) is the grid
function moved by int(s(r) / dtheta) whole steps with periodic wrap
(shift_function, :393-397), and the sigmoid derivatives are polynomials in
the sigmoid (sigmoid_nth_derivative, :212-243).
"""
from __future__ import annotations

import ctypes
import json
import math
import os

import numpy as np
from numpy.polynomial import Polynomial

from . import _lib

BZ, BR, BFI = 0, 1, 2


def calculate_s(r_axis, s0, alpha_deg):
    """Spiral azimuth s(r): s0 plus the running sum of tan(alpha) / r dr."""
    s = np.zeros_like(np.asarray(r_axis, dtype=np.float64))
    s[0] = s0
    ta = np.tan(np.deg2rad(alpha_deg))
    for i in range(1, len(r_axis)):
        s[i] = s[i - 1] + ta / r_axis[i] * (r_axis[i] - r_axis[i - 1])
    return s


def BmapParams_to_ConvexParams(Bmap):
    """Poles (x0, x1, W0, (|polarity|, sign)) in radians from the
    ``interval_1`` / ``positive_or_negative_1`` / ``fringe_width_1`` lists;
    zero-polarity intervals only advance the start angle."""
    out = []
    sector = 360.0 / Bmap["NSector"]
    start = 0.0
    for frac, pol, fringe in zip(Bmap["interval_1"], Bmap["positive_or_negative_1"], Bmap["fringe_width_1"]):
        if pol == 0:
            start += frac * sector
            continue
        end = start + frac * sector
        out.append((np.deg2rad(start), np.deg2rad(end), np.deg2rad((end - start) * fringe),
                    (abs(pol), 1 if pol > 0 else -1)))
        start = end
    return out


# ---------------------------------------------------------------------------
# operator algebra (host, exact integer coefficients)
# ---------------------------------------------------------------------------
def _add(dst, key, c):
    v = dst.get(key, 0) + c
    if v:
        dst[key] = v
    else:
        dst.pop(key, None)


def _op_dr(ops):
    """d/dr of sum c r^-p d^a_r d^b_theta."""
    out = {}
    for (a, b, p), c in ops.items():
        _add(out, (a + 1, b, p), c)
        if p:
            _add(out, (a, b, p + 1), -c * p)
    return out


def _op_dtheta(ops):
    return {(a, b + 1, p): c for (a, b, p), c in ops.items()}


def _op_laplacian(ops):
    out = {}
    for (a, b, p), c in _op_dr(_op_dr(ops)).items():
        _add(out, (a, b, p), c)
    for (a, b, p), c in _op_dr(ops).items():
        _add(out, (a, b, p + 1), c)
    for (a, b, p), c in _op_dtheta(_op_dtheta(ops)).items():
        _add(out, (a, b, p + 2), c)
    return out


def operator_terms(max_order):
    """{(component, n): {(a, b, p): c}} for Lap^n (Bz), d/dr Lap^n (Br) and
    d/dtheta Lap^n (Bfi), n = 0..max_order."""
    out = {}
    lap = {(0, 0, 0): 1}
    for n in range(max_order + 1):
        out[(BZ, n)] = dict(lap)
        out[(BR, n)] = _op_dr(lap)
        out[(BFI, n)] = _op_dtheta(lap)
        lap = _op_laplacian(lap)
    return out


def chain_rule_terms(max_i):
    """[i] -> {(g, (e_1, e_2, ...)): c}: d^i/dr^i G(theta - s(r)) =
    sum c G^(g) prod_j (s^(j))^(e_j)  (Faa di Bruno, built by repeated d/dr)."""
    table = [{(0, ()): 1}]
    for _ in range(max_i):
        nxt = {}
        for (g, e), c in table[-1].items():
            up = list(e) or [0]
            up[0] += 1
            _add(nxt, (g + 1, tuple(up)), -c)                # G' du/dr with du/dr = -s'
            for j, ej in enumerate(e):
                if ej:
                    dn = list(e) + ([0] if j + 1 == len(e) else [])
                    dn[j] -= 1
                    dn[j + 1] += 1
                    while dn and dn[-1] == 0:
                        dn.pop()
                    _add(nxt, (g, tuple(dn)), c * ej)        # d/dr (s^(j+1))^ej
        table.append(nxt)
    return table


# ---------------------------------------------------------------------------
# per-radius scalars
# ---------------------------------------------------------------------------
def radial_derivatives(coef_k, r, max_d):
    """[d] = d^d/dr^d A(r) for A = a r^k (``coef_k[0]`` a number) or
    A = sum_i c_i (r - x0)^i (``coef_k[0]`` the coefficients, highest power
    first; ``coef_k[1]`` = x0)."""
    a, k = coef_k
    r = np.asarray(r, dtype=np.float64)
    out = []
    for d in range(max_d + 1):
        if isinstance(a, (int, float, np.floating, np.integer)):
            f = a * np.prod(np.append(np.arange(k, k - d, -1), 1))
            out.append(f * r ** (k - d))
        else:
            acc = np.zeros_like(r)
            for i, ci in enumerate(list(a)[::-1]):
                if i >= d:
                    acc = acc + ci * math.factorial(i) / math.factorial(i - d) * (r - k) ** (i - d)
            out.append(acc)
    return np.array(out)


def taylor_weights(max_order, r_axis, s_poly, coef_k):
    """Weights [nr][3][max_order+1][2 max_order+2] (components Bz, Br, Bfi)."""
    r = np.asarray(r_axis, dtype=np.float64)
    n_orders, n_q = max_order + 1, 2 * max_order + 2
    max_a = 2 * max_order + 1
    A = radial_derivatives(coef_k, r, max_a)
    S = [s_poly.deriv(m=j)(r) for j in range(1, max_a + 1)]       # S[j-1] = s^(j)(r)
    chain = chain_rule_terms(max_a)
    # chain-rule factor per (i, g): sum c prod (s^(j))^e_j, as arrays over r
    chain_w = {}
    for i, terms in enumerate(chain):
        for (g, e), c in terms.items():
            f = np.full_like(r, float(c))
            for j, ej in enumerate(e):
                if ej:
                    f = f * S[j] ** ej
            chain_w[(i, g)] = chain_w.get((i, g), 0.0) + f
    W = np.zeros((r.size, 3, n_orders, n_q))
    for (comp, n), ops in operator_terms(max_order).items():
        pref = (-1) ** n / math.factorial(2 * n + (0 if comp == BZ else 1))
        acc = np.zeros((n_q, r.size))
        for (a, b, p), c in ops.items():
            rp = float(c) / r ** p
            for i in range(a + 1):
                lead = math.comb(a, i) * A[a - i] * rp
                for g in range(i + 1):
                    f = chain_w.get((i, g))
                    if f is not None:
                        acc[g + b] += lead * f
        if comp == BFI:
            acc = acc * (pref / r)
        else:
            acc = acc * pref
        W[:, comp, n, :] = acc.T
    return W


def sigmoid_polynomials(k, n_q):
    """[i][n_q+1]: the i-th derivative of 1/(1+exp(-k(x-x0))) as a
    polynomial in the sigmoid, highest power first, left-padded; P_1 =
    k s - k s^2, P_(i+1) = P_i' P_1; row 0 is s itself."""
    out = np.zeros((n_q, n_q + 1))
    out[0, -2] = 1.0
    p1 = np.array([-k, k, 0.0])
    p = p1
    for i in range(1, n_q):
        if i > 1:
            deg = len(p) - 1
            p = np.convolve(np.array([p[t] * (deg - t) for t in range(deg)]), p1)
        out[i, n_q + 1 - len(p):] = p
    return out


def _pole_table(convex_paras):
    """(x0, x1, k = 16 / W0, height) rows from reference-style poles
    (x0, x1, W0, (|polarity|, sign)) or 5-column rows (x0, x1, W0, |pol|, sign)."""
    rows = []
    for pr in convex_paras:
        if len(pr) == 4:
            x0, x1, w0, (mag, sgn) = pr
        else:
            x0, x1, w0, mag, sgn = pr
        rows.append((x0, x1, 16.0 / w0, sgn * abs(mag)))
    return np.array(rows, dtype=np.float64)


class TaylorMapGenerator:
    """Device maps for one lattice: ``run(layout, dtype)`` returns a torch
    tensor on the context's device -- ``"stacks"``: (3, max_order+1, nr,
    n_phi) with components (Bz, Br, Bfi); ``"tracking"``: (n_phi, nr,
    3 max_order), orders 1.. with (Bz, Bfi, Br) per order, the table the
    pushers read."""

    def __init__(self, max_order, r_axis, fi_axis_rad, convex_paras, s_poly, coef_k):
        if not 0 <= max_order < _lib.FFA_TAYLOR_MAX_ORDERS:
            raise ValueError(f"max_order must be in [0, {_lib.FFA_TAYLOR_MAX_ORDERS - 1}]")
        self.max_order = int(max_order)
        self.r_axis = np.asarray(r_axis, dtype=np.float64)
        self.fi_axis = np.asarray(fi_axis_rad, dtype=np.float64)
        self.n_q = 2 * self.max_order + 2
        self.poles = _pole_table(convex_paras)
        if not 1 <= len(self.poles) <= _lib.FFA_TAYLOR_MAX_POLES:
            raise ValueError(f"between 1 and {_lib.FFA_TAYLOR_MAX_POLES} poles are supported")
        self.weights = taylor_weights(self.max_order, self.r_axis, s_poly, coef_k)
        step = self.fi_axis[1] - self.fi_axis[0]
        sh = np.array([int(v / step) for v in s_poly(self.r_axis)], dtype=np.int64)
        self.shift = np.mod(sh, self.fi_axis.size).astype(np.int32)
        self.sig_poly = np.stack([np.stack([sigmoid_polynomials(k, self.n_q), sigmoid_polynomials(-k, self.n_q)])
                                  for k in self.poles[:, 2]])

    def output_shape(self, layout):
        nr, nphi, no = self.r_axis.size, self.fi_axis.size, self.max_order + 1
        return (3, no, nr, nphi) if layout == "stacks" else (nphi, nr, 3 * (no - 1))

    def run(self, layout="stacks", dtype=None, ctx=None, out=None):
        """Launches the generator on the current stream and returns the map
        tensor (``out`` when given: a device tensor of the right shape and
        type).  The inputs are uploaded once per device and kept."""
        import torch
        if ctx is None:
            ctx = self._ctx = getattr(self, "_ctx", None) or _lib.Context(torch.cuda.current_device(),
                                                                          torch.cuda.current_stream().cuda_stream)
        dev = torch.device("cuda", torch.cuda.current_device())
        dtype = dtype or torch.float64
        if dtype not in (torch.float64, torch.float32):
            raise ValueError("dtype must be float64 or float32")
        if layout not in ("stacks", "tracking"):
            raise ValueError("layout must be 'stacks' or 'tracking'")
        nr, nphi, no = self.r_axis.size, self.fi_axis.size, self.max_order + 1
        if layout == "tracking" and no < 2:
            raise ValueError("the tracking layout needs max_order >= 1")

        def up(a, dt=torch.float64):
            return torch.from_numpy(np.ascontiguousarray(a)).to(device=dev, dtype=dt)
        keep = getattr(self, "_dev", None)
        if keep is None or keep["theta"].device != dev:
            keep = self._dev = dict(theta=up(self.fi_axis), poles=up(self.poles), sig=up(self.sig_poly),
                                    w=up(self.weights), shift=up(self.shift, torch.int32),
                                    table=torch.empty((self.n_q, nphi), dtype=torch.float64, device=dev))
        shape = self.output_shape(layout)
        if out is None:
            out = torch.empty(shape, dtype=dtype, device=dev)
        elif tuple(out.shape) != shape or out.dtype != dtype or out.device != dev or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous {dtype} tensor of shape {shape} on {dev}")
        g = _lib.FfaTaylorGen(n_orders=no, n_q=self.n_q, n_poles=len(self.poles),
                              layout=0 if layout == "stacks" else 1, out_f32=int(dtype == torch.float32), pad=0,
                              nr=nr, n_phi=nphi, theta=keep["theta"].data_ptr(), poles=keep["poles"].data_ptr(),
                              sig_poly=keep["sig"].data_ptr(), weights=keep["w"].data_ptr(),
                              shift=keep["shift"].data_ptr(), flutter=keep["table"].data_ptr(),
                              out=out.data_ptr())
        ctx.check(ctx.lib.ffa_taylor_maps(ctx.h, ctypes.byref(g)), "ffa_taylor_maps")
        return out


def Generate_Bmaps_MultiProcess(MaxOrder, r_axis, fi_axis_rad, convex_paras, s_values_FitCoef_rad, BzR_coef_k):
    """Lists (Br, Bz, Bfi) of host arrays (len(r_axis), len(fi_axis_rad)),
    orders 0..MaxOrder, computed on the current GPU."""
    import torch
    gen = TaylorMapGenerator(MaxOrder, r_axis, fi_axis_rad, convex_paras, s_values_FitCoef_rad, BzR_coef_k)
    st = gen.run("stacks", torch.float64).cpu().numpy()
    return ([st[BR, n] for n in range(MaxOrder + 1)], [st[BZ, n] for n in range(MaxOrder + 1)],
            [st[BFI, n] for n in range(MaxOrder + 1)])


def grid_from_config(Bmap):
    """(r_axis, fi_axis) with the int(x)+1 point counts of main_Bmap."""
    rmin, rmax, rstep = Bmap["rmin_max_step_m"]
    n_r = int((rmax - rmin) / rstep) + 1
    n_fi = int((np.pi * 2 / Bmap["NSector"]) / Bmap["theta_step_rad"]) + 1
    return np.linspace(rmin, rmax, n_r), np.linspace(0, np.pi * 2 / Bmap["NSector"], n_fi)


def main_Bmap(json_path, expand_order=None):
    """Writes the Taylor maps of the lattice in ``json_path`` (a
    config_Bmap.json with ``Bmap`` and ``config`` = {folder, expand_order})
    into ``config.folder`` (relative to the JSON file) and returns that
    folder."""
    json_path = os.path.abspath(json_path)
    with open(json_path) as f:
        cfg = json.load(f)
    bm = cfg["Bmap"]
    folder = cfg["config"]["folder"]
    order = cfg["config"]["expand_order"] if expand_order is None else int(expand_order)
    out_dir = folder if os.path.isabs(folder) else os.path.join(os.path.dirname(json_path), folder)
    os.makedirs(out_dir, exist_ok=True)
    r_axis, fi_axis = grid_from_config(bm)
    s_fit = Polynomial.fit(r_axis, calculate_s(r_axis, 0.0, bm["SpiralAngle_deg"]), 5)
    convex = BmapParams_to_ConvexParams(bm)
    br, bz, bfi = Generate_Bmaps_MultiProcess(order, r_axis, fi_axis, convex, s_fit, tuple(bm["coefficients_tupel"]))
    for n in range(order + 1):
        np.savez(os.path.join(out_dir, f"Br_order_{n}.npz"), Br_Taylor_coeff=br[n])
        np.savez(os.path.join(out_dir, f"Bz_order_{n}.npz"), Bz_Taylor_coeff=bz[n])
        np.savez(os.path.join(out_dir, f"Bfi_order_{n}.npz"), Bfi_Taylor_coeff=bfi[n])
    # Bmean: azimuthal mean of the order-0 field per radius (UpdataConfig)
    cfg["Bmap"]["Bmean"] = bz[0].mean(axis=1).tolist()
    cfg["Bmap"]["r_axis"] = r_axis.tolist()
    with open(os.path.join(out_dir, "config_Bmap.json"), "w") as f:
        json.dump(cfg, f, indent=4)
    return out_dir
