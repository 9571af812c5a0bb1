"""Module-level stages of the 2.5D space-charge solve, on the GPU.

The names and positional signatures are those of the reference's
``src/FFAG_SC_new.py``, which ``FFAG_Bunch.Build_2_5D_SliceGrid_sbs``
(src/FFAG_Bunch.py:1157-1352) and the SC benchmarks call stage by stage.
Each function runs its stage in a HIP kernel of libffa_hip.so (the
``ffa_sc_*`` stage entry points of include/ffa_hip.h) and keeps the
reference's conventions: the same array layouts, outputs written into the
caller's buffers where the reference writes in place, the same return
values.  Arguments may be numpy arrays (copied to the GPU and the results
copied back) or CUDA torch tensors (used in place, nothing copied).

The tracking loop does not call these: it runs the fused solve
(``ffa_sc_solve``), whose stages are the same arithmetic.  Differences from
the reference, by design: sums (sum of gamma, moments) are fixed-order tree
reductions, reproducible run to run but not the reference's sequential
rounding; the deposit rounds each particle's CIC fractions to fixed point
and sums integers (bit-reproducible; the reference accumulates float32 in
particle order); the Poisson solve is the LDS Stockham FFT in complex64.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

_state = {}


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.FfaError("FFAG_SC_new runs on the GPU: no HIP device is available")
    return torch


def _ctx():
    torch = _torch()
    dev = torch.cuda.current_device()
    stream = torch.cuda.current_stream().cuda_stream
    c = _state.get("ctx")
    if c is None or _state.get("dev") != dev:
        c = _lib.Context(dev, stream)
        _state.update(ctx=c, dev=dev, ws=None)
    else:
        c.check(c.lib.ffa_set_stream(c.h, _lib._vp(stream)), "ffa_set_stream")
    return c


def _workspace(ctx, n, ns, nz, nx):
    torch = _torch()
    nbytes = ctypes.c_size_t()
    ctx.check(ctx.lib.ffa_sc_stage_workspace_size(int(n), int(ns), int(nz), int(nx), ctypes.byref(nbytes)),
              "ffa_sc_stage_workspace_size")
    ws = _state.get("ws")
    if ws is None or ws.numel() < nbytes.value:
        ws = torch.empty(max(nbytes.value, 256), dtype=torch.uint8, device="cuda")
        _state["ws"] = ws
    return ws, ws.numel()


def _is_dev(a):
    return hasattr(a, "is_cuda") and a.is_cuda


def _dev(a, dtype):
    """Device tensor of ``a`` with the given torch dtype (shared when ``a``
    already is a contiguous CUDA tensor of that type)."""
    torch = _torch()
    if _is_dev(a):
        t = a if a.dtype == dtype else a.to(dtype)
        return t.contiguous()
    return torch.as_tensor(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def _p(t):
    return _lib._vp(t.data_ptr()) if t is not None else None


def _back(dst, src):
    """Write a device result into the caller's buffer (numpy or tensor)."""
    if _is_dev(dst):
        dst.copy_(src.reshape(dst.shape).to(dst.dtype))
    else:
        np.copyto(dst, src.reshape(dst.shape).cpu().numpy().astype(dst.dtype, copy=False))
    return dst


def _host_or_dev(like, t):
    return t if _is_dev(like) else t.cpu().numpy()


# ---------------------------------------------------------------------------
# GetLocalCoordinates_2p5 / _rk (src/FFAG_SC_new.py:74-148)
# ---------------------------------------------------------------------------
def _local_coords(LocalBunch, Step_Survive, malloc_r, malloc_fi, malloc_z, malloc_gamma, malloc_mean_gamma,
                  malloc_fiMod, Malloc_SC_surviveIDX, speed_c, rk):
    torch = _torch()
    ctx = _ctx()
    M = _dev(LocalBunch, torch.float64)
    n = M.shape[0]
    surv = _dev(Step_Survive, torch.uint8)
    out = torch.empty(6 * max(n, 1), dtype=torch.float64, device="cuda")
    cs = torch.zeros(2, dtype=torch.float64, device="cuda")
    ws, nb = _workspace(ctx, n, 1, 1, 1)
    ctx.check(ctx.lib.ffa_sc_local_coords(ctx.h, _p(M), _p(surv), n, float(speed_c), 1 if rk else 0, _p(out), _p(cs),
                                          _p(ws), nb), "ffa_sc_local_coords")
    count, sum_gamma = cs.cpu().tolist()
    j = int(count)
    planes = out[: 6 * n].view(6, n)[:, :j] if n else out[:0].view(6, 0)
    for dst, k in ((malloc_r, 0), (malloc_fi, 1), (malloc_z, 2), (malloc_gamma, 3), (malloc_fiMod, 4),
                   (Malloc_SC_surviveIDX, 5)):
        _back(dst[:j], planes[k])
    malloc_mean_gamma[:j] = 1.0
    return (malloc_r[:j], malloc_fi[:j], malloc_z[:j], malloc_gamma[:j], malloc_mean_gamma[:j], malloc_fiMod[:j],
            Malloc_SC_surviveIDX[:j], j, sum_gamma)


def GetLocalCoordinates_2p5(LocalBunch, Step_Survive, malloc_r, malloc_fi, malloc_z, malloc_gamma, malloc_mean_gamma,
                            malloc_fiMod, Malloc_SC_surviveIDX, speed_c):
    """Survivors' (r, fi, z, gamma, fi mod 2 pi, row) in row order, gamma
    from the Boris state (src/FFAG_SC_new.py:74-108); ffa_sc_local_coords."""
    return _local_coords(LocalBunch, Step_Survive, malloc_r, malloc_fi, malloc_z, malloc_gamma, malloc_mean_gamma,
                         malloc_fiMod, Malloc_SC_surviveIDX, speed_c, False)


def GetLocalCoordinates_2p5_rk(LocalBunch, Step_Survive, malloc_r, malloc_fi, malloc_z, malloc_gamma,
                               malloc_mean_gamma, malloc_fiMod, Malloc_SC_surviveIDX, speed_c):
    """As GetLocalCoordinates_2p5 with gamma = E_total / E0 from the RK4
    state (src/FFAG_SC_new.py:111-148)."""
    return _local_coords(LocalBunch, Step_Survive, malloc_r, malloc_fi, malloc_z, malloc_gamma, malloc_mean_gamma,
                         malloc_fiMod, Malloc_SC_surviveIDX, speed_c, True)


# ---------------------------------------------------------------------------
# Covering arc (src/FFAG_SC_new.py:534-606)
# ---------------------------------------------------------------------------
def divide_arc2(points_fi):
    """(arc_start, arc_end, arc_len) of the smallest cover among the four
    shifted ones (src/FFAG_SC_new.py:534-566); ffa_sc_divide_arc."""
    torch = _torch()
    ctx = _ctx()
    fi = _dev(points_fi, torch.float64)
    out = torch.zeros(3, dtype=torch.float64, device="cuda")
    ws, nb = _workspace(ctx, 0, 1, 1, 1)
    ctx.check(ctx.lib.ffa_sc_divide_arc(ctx.h, _p(fi), fi.numel(), _p(out), _p(ws), nb), "ffa_sc_divide_arc")
    s, e, span = out.cpu().tolist()
    return s, e, span


def expand_arc(start, end):
    """An arc as a non-decreasing interval (src/FFAG_SC_new.py:570-574)."""
    if end < start:
        end += 2 * np.pi
    return start, end


def merge_multiple_arcs(arcs):
    """Smallest arc covering the ranks' arcs (src/FFAG_SC_new.py:577-606):
    each later arc is moved by the whole turn that brings its centre nearest
    the first arc's centre, then the union's ends are taken mod 2 pi.  A few
    arcs per solve (one per rank): host arithmetic."""
    expanded = [expand_arc(s, e) for s, e in arcs]
    if not expanded:
        return 0.0, np.pi * 2
    centre = (expanded[0][0] + expanded[0][1]) / 2
    kept = [expanded[0]]
    for s, e in expanded[1:]:
        options = [(s, e), (s + 2 * np.pi, e + 2 * np.pi), (s - 2 * np.pi, e - 2 * np.pi)]
        kept.append(min(options, key=lambda a: abs((a[0] + a[1]) / 2 - centre)))
    return min(a[0] for a in kept) % (2 * np.pi), max(a[1] for a in kept) % (2 * np.pi)


# ---------------------------------------------------------------------------
# Slicing and moments (src/FFAG_SC_new.py:449-531)
# ---------------------------------------------------------------------------
def ComputeSliceBoundsAndCoordsSBS_merged(coords_r, coords_z, coords_fi_mod, coords_r0, coords_x, slice_fi_step,
                                          fi_s_global, segments, r2d, z2d, x2d, slice_indices):
    """x = r - r0, the slice of every particle, the per-slice lists r2d,
    z2d, x2d (segments rows, particle order) and the moment sums;
    ffa_sc_slices."""
    torch = _torch()
    ctx = _ctx()
    r = _dev(coords_r, torch.float64)
    n = r.numel()
    z = _dev(coords_z, torch.float64)
    fm = _dev(coords_fi_mod, torch.float64)
    r0 = _dev(coords_r0, torch.float64)
    ns = int(segments)
    stride = int(r2d.shape[1])
    for a in (r2d, z2d, x2d):
        if a.ndim != 2 or a.shape[0] < ns or a.shape[1] != stride:
            raise ValueError("r2d, z2d, x2d must be (segments, Nmax) arrays of one shape")
    x = torch.empty(max(n, 1), dtype=torch.float64, device="cuda")
    sl = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    counts = torch.empty(ns, dtype=torch.int32, device="cuda")
    sums = torch.zeros(4, dtype=torch.float64, device="cuda")
    lists = [_dev(a, torch.float64) if _is_dev(a) else torch.zeros(ns * stride, dtype=torch.float64, device="cuda")
             for a in (r2d, z2d, x2d)]
    ws, nb = _workspace(ctx, n, ns, 1, 1)
    ctx.check(ctx.lib.ffa_sc_slices(ctx.h, _p(r), _p(z), _p(fm), _p(r0), n, float(slice_fi_step), float(fi_s_global),
                                    ns, _p(x), _p(lists[0]), _p(lists[1]), _p(lists[2]), stride, _p(counts), _p(sl),
                                    _p(sums), _p(ws), nb), "ffa_sc_slices")
    _back(coords_x[:n], x[:n])
    _back(slice_indices[:n], sl[:n])
    if not _is_dev(r2d):
        # only the filled part of each row is written, as the reference does
        c = counts.cpu().numpy()
        for dst, src in zip((r2d, z2d, x2d), lists):
            h = src.view(ns, stride).cpu().numpy()
            for s in np.flatnonzero(c):
                dst[s, :c[s]] = h[s, :c[s]]
    sx, sz, sx2, sz2 = sums.cpu().tolist()
    cnt = _host_or_dev(coords_r, counts)
    return coords_x[:n], r2d, z2d, x2d, cnt, slice_indices, sx, sz, sx2, sz2


# ---------------------------------------------------------------------------
# Deposit (src/FFAG_SC_new.py:244-312)
# ---------------------------------------------------------------------------
def fixed_point_bits(n_particles: int) -> int:
    """Fractional bits per axis of the fixed-point CIC weights: at most 20,
    and n * 2^(2 bits) < 2^63 so no int64 cell can overflow."""
    b = 20
    while b > 1 and n_particles * (1 << (2 * b)) >= (1 << 63):
        b -= 1
    return b


def DistributeChargeSBS_Njit(z_slices, x_slices, count_slices, q_macro, BunchLength, xmin, xmax, zmin, zmax, nx, nz,
                             ns):
    """Slice-by-slice CIC charge density (ns, nz, nx) float32 with the
    normalisation q_macro ns / (BunchLength dx dz); zeros for a degenerate
    box (src/FFAG_SC_new.py:244-312); ffa_sc_deposit (fixed point)."""
    torch = _torch()
    ctx = _ctx()
    zs = _dev(z_slices, torch.float64)
    xs = _dev(x_slices, torch.float64)
    cnt = _dev(count_slices, torch.int32)
    stride = int(zs.shape[1]) if zs.dim() == 2 else 0
    total = int(cnt.sum().item())
    rho = torch.empty((int(ns), int(nz), int(nx)), dtype=torch.float32, device="cuda")
    ws, nb = _workspace(ctx, 0, ns, nz, nx)
    ctx.check(ctx.lib.ffa_sc_deposit(ctx.h, _p(zs), _p(xs), stride, _p(cnt), int(ns), int(nz), int(nx), float(q_macro),
                                     float(BunchLength), float(xmin), float(xmax), float(zmin), float(zmax),
                                     fixed_point_bits(max(total, 1)), _p(rho), _p(ws), nb), "ffa_sc_deposit")
    return _host_or_dev(z_slices, rho)


# ---------------------------------------------------------------------------
# Green's function, Poisson solve, field (src/FFAG_SC_new.py:672-760)
# ---------------------------------------------------------------------------
def make_fft_kernel_open_numba(nx, nz, dx, dz):
    """Periodic Green's function 1/(eps0 k^2), (nz, nx) float32, zero k = 0
    mode, in the reference's float32 arithmetic (src/FFAG_SC_new.py:731-760);
    ffa_sc_green."""
    torch = _torch()
    ctx = _ctx()
    out = torch.empty((int(nz), int(nx)), dtype=torch.float32, device="cuda")
    ws, nb = _workspace(ctx, 0, 2, nz, nx)
    ctx.check(ctx.lib.ffa_sc_green(ctx.h, int(nx), int(nz), float(dx), float(dz), _p(out), _p(ws), nb), "ffa_sc_green")
    return out.cpu().numpy()


def poisson_solve_slices(rho, dx, dz):
    """phi_k = Re IFFT2(FFT2(rho_k) G) / (nz nx) for every slice k of
    rho (ns, nz, nx): the per-slice pyFFTW loop of Build_2_5D_SliceGrid_sbs
    (src/FFAG_Bunch.py:1330-1338) on the LDS Stockham FFT; ffa_sc_poisson."""
    torch = _torch()
    ctx = _ctx()
    r = _dev(rho, torch.float32)
    ns, nz, nx = (int(v) for v in r.shape)
    phi = torch.empty_like(r)
    ws, nb = _workspace(ctx, 0, ns, nz, nx)
    ctx.check(ctx.lib.ffa_sc_poisson(ctx.h, _p(r), ns, nz, nx, float(dx), float(dz), _p(phi), _p(ws), nb),
              "ffa_sc_poisson")
    return _host_or_dev(rho, phi)


def compute_field_from_potential_numba_SBS(phi, k, dr, dz, BunchLength, Ex_cube, Ez_cube, mean_gamma_global):
    """Ex_cube[k], Ez_cube[k] = -grad(phi) / gamma^2 with the reference's
    edge stencils, zeros for a degenerate grid (src/FFAG_SC_new.py:672-728);
    ffa_sc_field."""
    torch = _torch()
    ctx = _ctx()
    p = _dev(phi, torch.float32)
    nz, nx = (int(v) for v in p.shape[-2:])
    ex = torch.empty((1, nz, nx), dtype=torch.float32, device="cuda")
    ez = torch.empty_like(ex)
    ws, nb = _workspace(ctx, 0, 1, 1, 1)
    ctx.check(ctx.lib.ffa_sc_field(ctx.h, _p(p), 1, nz, nx, float(dr), float(dz), float(BunchLength),
                                   float(mean_gamma_global), _p(ex), _p(ez), _p(ws), nb), "ffa_sc_field")
    _back(Ex_cube[k], ex[0])
    _back(Ez_cube[k], ez[0])


def total_field_1d_numba(z, q):
    """Longitudinal Coulomb field of the slice charges q at positions z
    (src/FFAG_SC_new.py:763-782); ffa_sc_total_field_1d."""
    torch = _torch()
    ctx = _ctx()
    s = _dev(z, torch.float64)
    qq = _dev(q, torch.float64)
    E = torch.zeros(max(s.numel(), 1), dtype=torch.float64, device="cuda")
    ctx.check(ctx.lib.ffa_sc_total_field_1d(ctx.h, _p(s), _p(qq), s.numel(), _p(E)), "ffa_sc_total_field_1d")
    return _host_or_dev(z, E[: s.numel()])


# ---------------------------------------------------------------------------
# Gather (src/FFAG_SC_new.py:870-972)
# ---------------------------------------------------------------------------
def interpolate_fields_to_particles_SBS_numba(Ex_cube, Ez_cube, Ef, coords_x, coords_z, coords_pr0, slice_indices,
                                              coords_survive_idx, x0, z0, dx, dz, BunchLength, Ex_p, Ez_p, Ef_p,
                                              mat_in):
    """Bilinear Ex, Ez of each particle's slice, projected with pr0, into
    Ex_p, Ez_p, Ef_p and columns 10-12 of mat_in's rows coords_survive_idx
    (src/FFAG_SC_new.py:870-972); ffa_sc_interpolate.  Ef is accepted for the
    signature and unused, as in the reference."""
    torch = _torch()
    ctx = _ctx()
    ex = _dev(Ex_cube, torch.float32)
    ez = _dev(Ez_cube, torch.float32)
    ns, nz, nx = (int(v) for v in ex.shape)
    x = _dev(coords_x, torch.float64)
    n = x.numel()
    z = _dev(coords_z, torch.float64)
    pr0 = _dev(coords_pr0, torch.float64)
    sl = _dev(slice_indices, torch.int32)
    rows = _dev(np.asarray(coords_survive_idx).astype(np.int64) if not _is_dev(coords_survive_idx)
                else coords_survive_idx, torch.int64)
    M = _dev(mat_in, torch.float64) if mat_in is not None else None
    outs = [torch.empty(max(n, 1), dtype=torch.float64, device="cuda") for _ in range(3)]
    ctx.check(ctx.lib.ffa_sc_interpolate(ctx.h, _p(ex), _p(ez), ns, nz, nx, _p(x), _p(z), _p(pr0), _p(sl), _p(rows), n,
                                         float(x0), float(z0), float(dx), float(dz), float(BunchLength), _p(outs[0]),
                                         _p(outs[1]), _p(outs[2]), _p(M), M.shape[0] if M is not None else 0),
              "ffa_sc_interpolate")
    for dst, src in zip((Ex_p, Ez_p, Ef_p), outs):
        _back(dst[:n], src[:n])
    if M is not None and not _is_dev(mat_in):
        idx = np.asarray(coords_survive_idx).astype(np.int64)[:n]
        mat_in[idx, 10:13] = M[torch.as_tensor(idx, device="cuda")][:, 10:13].cpu().numpy()
    elif M is not None and M.data_ptr() != mat_in.data_ptr():
        mat_in.copy_(M)
    return Ex_p, Ez_p, Ef_p
