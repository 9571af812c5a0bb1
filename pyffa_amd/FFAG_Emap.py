"""RF frequency curves and spiral-gap field maps (the Emap builder).

Same entry points and files as examples/main_Emap.py:

* ``build_frequency_curve_from_config(config, Emap, machine)`` (:448-862):
  steps the synchronous particle gap by gap (energy gain V0 sin(phi_s),
  gap-to-gap time from the revolution frequency f_rev(E) of SEO_ini.txt,
  corrected for the spiral gaps' azimuth change), derives the RF frequency
  h / (Ngap dt) + d(phi_s) / (2 pi dt), resamples (t, f, V0, E) on a
  uniform time grid and writes the frequency-curve file with its header;
* ``build_phi_gap_interpolator_from_SEO`` (:149-264): where the spiral
  centre line crosses each closed orbit of SEO_r.txt;
* ``build_spiral_map_from_config(config, Emap)`` (:865-985): the Er, Ez/z,
  Ephi maps of a Gaussian gap along the spiral line, computed on the GPU
  (``ffa_spiral_gap_maps``: a Newton nearest-point search per map node);
* ``main_Emap(json_path)`` (:1147-1176) runs both from a config_Emap.json.

The frequency-curve stepping is a sequential recurrence (each gap's time
depends on the energy after the previous one) and stays on the host; the map
has one independent Newton search per node and runs on the GPU.
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from . import _lib


def wrap_0_2pi(x):
    return np.mod(x, 2.0 * np.pi)


def wrap_pm_pi(x):
    twopi = 2.0 * np.pi
    y = (x + np.pi) % twopi - np.pi
    return np.where(y <= -np.pi, y + twopi, y)


def phi_spiral_of_r(r, r_ref, s_ref, tan_a):
    return s_ref + tan_a * np.log(np.asarray(r, dtype=np.float64) / r_ref)


def _linear_extrap(x, y):
    """Linear interpolation that continues the end segments outside."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)

    def f(q):
        q = np.asarray(q, dtype=np.float64)
        i = np.clip(np.searchsorted(x, q, side="right") - 1, 0, x.size - 2)
        return y[i] + (q - x[i]) / (x[i + 1] - x[i]) * (y[i + 1] - y[i])
    return f


def interp_periodic_phi(phi_query, phi_base, val_base):
    """Periodic linear interpolation of a function of the azimuth (the base
    points wrapped, sorted, de-duplicated and tiled over three turns)."""
    pw = wrap_0_2pi(np.asarray(phi_base, dtype=np.float64))
    order = np.argsort(pw)
    ps, vs = pw[order], np.asarray(val_base, dtype=np.float64)[order]
    keep = np.ones(ps.size, bool)
    keep[1:] = np.diff(ps) > 1e-14
    ps, vs = ps[keep], vs[keep]
    if ps.size < 2:
        raise ValueError("Too few unique phi points after periodic deduplication.")
    two = 2.0 * np.pi
    return np.interp(phi_query, np.concatenate([ps - two, ps, ps + two]), np.concatenate([vs, vs, vs]))


def find_intersections_phi_theta(phi, r_co, phi_curve, rmin=None, rmax=None, atol=1e-12):
    """Azimuths where the orbit sample (phi, r_co) meets the spiral line
    phi_curve(r_co): sign changes of wrap(phi_curve - phi), linearly refined,
    skipping jumps across +-pi."""
    g = wrap_pm_pi(phi_curve - phi)
    out = []
    for k in range(len(g) - 1):
        g0, g1 = g[k], g[k + 1]
        if abs(g1 - g0) > np.pi:
            continue
        if abs(g0) < atol:
            if (rmin is None or r_co[k] >= rmin) and (rmax is None or r_co[k] <= rmax):
                out.append((phi[k], r_co[k]))
            continue
        if g0 * g1 < 0.0 or abs(g1) < atol:
            den = g1 - g0
            if abs(den) < 1e-15:
                continue
            t = -g0 / den
            ps, rs = phi[k] + t * (phi[k + 1] - phi[k]), r_co[k] + t * (r_co[k + 1] - r_co[k])
            if (rmin is None or rs >= rmin) and (rmax is None or rs <= rmax):
                out.append((ps, rs))
    return out


def build_phi_gap_interpolator_from_SEO(SEO_r_path, alpha_deg, r_ref, s_ref_deg, rmin_valid, rmax_valid, Nphi,
                                        interp_kind="linear", verbose=False):
    """(phi_gap(Ek) callable, Ek nodes, unwrapped gap azimuths, gap radii)."""
    data = np.loadtxt(SEO_r_path, skiprows=1)
    phi_seo, ek, r_mat = data[0, 1:], data[1:, 0], data[1:, 1:]
    order = np.argsort(ek)
    ek, r_mat = ek[order], r_mat[order]
    tan_a, s_ref = np.tan(np.deg2rad(alpha_deg)), np.deg2rad(s_ref_deg)
    if r_ref <= 0.0 or rmin_valid <= 0.0 or rmax_valid <= rmin_valid or int(Nphi) < 2:
        raise RuntimeError("invalid spiral or radius parameters")
    lo, hi = sorted((s_ref + tan_a * np.log(rmin_valid / r_ref), s_ref + tan_a * np.log(rmax_valid / r_ref)))
    pad = max((hi - lo) * 0.02, 1e-6)
    grid = np.linspace(lo - pad, hi + pad, int(Nphi))
    phi_pick, r_pick = np.full(ek.size, np.nan), np.full(ek.size, np.nan)
    for i in range(ek.size):
        r_co = interp_periodic_phi(grid, phi_seo, r_mat[i])
        ok = (r_co >= rmin_valid) & (r_co <= rmax_valid)
        if np.count_nonzero(ok) < 2:
            continue
        pts = find_intersections_phi_theta(grid[ok], r_co[ok], phi_spiral_of_r(r_co[ok], r_ref, s_ref, tan_a),
                                           rmin=rmin_valid, rmax=rmax_valid)
        if len(pts) > 1:
            raise RuntimeError(f"Ek={ek[i]:.6g} MeV has {len(pts)} intersections in valid range; expected exactly 1.")
        if pts:
            phi_pick[i], r_pick[i] = pts[0]
    good = ~np.isnan(phi_pick)
    if good.sum() < 2:
        raise RuntimeError("Too few valid spiral-gap intersections. Please check spiral parameters and r-range.")
    return _linear_extrap(ek[good], phi_pick[good]), ek[good], phi_pick[good], r_pick[good]


def build_phi_axis_from_spiral(r_min, r_max, r_ref, s_ref, alpha_rad, nphi_full, pad_deg, min_nphi):
    """(phi mod 2 pi, continuous phi) covering the spiral over [r_min, r_max]
    plus pad_deg on both sides, at about 2 pi / nphi_full spacing."""
    a = np.tan(alpha_rad)
    p1, p2 = s_ref + a * np.log(r_min / r_ref), s_ref + a * np.log(r_max / r_ref)
    lo, hi = min(p1, p2) - np.deg2rad(pad_deg), max(p1, p2) + np.deg2rad(pad_deg)
    nphi = max(int(np.ceil((hi - lo) / (2.0 * np.pi / nphi_full))), int(min_nphi))
    cont = np.linspace(lo, hi, nphi, endpoint=False)
    return np.mod(cont, 2.0 * np.pi), cont


def spiral_gap_maps(r_axis, phi_axis, r_ref, s_ref, alpha_rad, r_min, r_max, U, FWHM, n_iter=10, normal_side=1):
    """(Er, Ephi, Ez/z, distance, end flag) [nr][nphi] computed on the GPU."""
    import torch
    if not torch.cuda.is_available():
        raise _lib.FfaError("the spiral gap maps are computed on the GPU")
    dev = torch.device("cuda", torch.cuda.current_device())
    ctx = _lib.Context(dev.index, torch.cuda.current_stream().cuda_stream)
    r = torch.from_numpy(np.ascontiguousarray(r_axis, dtype=np.float64)).to(dev)
    p = torch.from_numpy(np.ascontiguousarray(phi_axis, dtype=np.float64)).to(dev)
    nr, nphi = r.numel(), p.numel()
    out = {k: torch.empty((nr, nphi), dtype=torch.float64, device=dev) for k in ("er", "ephi", "ez", "dist")}
    flag = torch.empty((nr, nphi), dtype=torch.int32, device=dev)
    g = _lib.FfaSpiralGap(r_axis=r.data_ptr(), phi_axis=p.data_ptr(), nr=nr, nphi=nphi,
                          dr=float(r_axis[1] - r_axis[0]), dphi=float(phi_axis[1] - phi_axis[0]), r_ref=float(r_ref),
                          s_ref=float(s_ref), a=float(np.tan(alpha_rad)), r_min=float(r_min), r_max=float(r_max),
                          U=float(U), delta=float(FWHM) / 2.35482, side=1.0 if normal_side >= 0 else -1.0,
                          n_iter=int(n_iter), pad=0, er=out["er"].data_ptr(), ephi=out["ephi"].data_ptr(),
                          ez_over_z=out["ez"].data_ptr(), dist=out["dist"].data_ptr(), flag=flag.data_ptr())
    ctx.check(ctx.lib.ffa_spiral_gap_maps(ctx.h, ctypes.byref(g)), "ffa_spiral_gap_maps")
    return (out["er"].cpu().numpy(), out["ephi"].cpu().numpy(), out["ez"].cpu().numpy(), out["dist"].cpu().numpy(),
            flag.cpu().numpy())


def build_spiral_map_from_config(config, Emap):
    """Writes Er / Ez / Ephi map files (first row the continuous azimuth,
    first column the radius, the saved radius range only)."""
    alpha = np.deg2rad(float(Emap["spiral_alpha_deg"]))
    r_min, r_max = float(Emap["spiral_rmin_calc"]), float(Emap["spiral_rmax_calc"])
    r_lo, r_hi = float(Emap["spiral_rmin_save"]), float(Emap["spiral_rmax_save"])
    r_ref, s_ref = float(Emap["spiral_r_ref"]), np.deg2rad(float(Emap["spiral_s_ref_deg"]))
    r_axis = np.linspace(r_min, r_max, int(Emap["spiral_nr"]))
    _, cont = build_phi_axis_from_spiral(r_min, r_max, r_ref, s_ref, alpha, int(Emap["spiral_nphi_full"]),
                                         float(Emap["spiral_pad_deg"]), int(Emap["spiral_min_nphi"]))
    er, ef, ez, _, _ = spiral_gap_maps(r_axis, cont, r_ref, s_ref, alpha, r_min, r_max, float(Emap["spiral_U"]),
                                       float(Emap["spiral_FWHM"]), int(Emap["spiral_newton_iter"]),
                                       int(Emap["spiral_normal_side"]))
    keep = (r_axis >= r_lo) & (r_axis <= r_hi)
    os.makedirs(Emap["emap_save_dir"], exist_ok=True)
    paths = []
    for key, tab in (("Ez_map_file", ez), ("Er_map_file", er), ("Ephi_map_file", ef)):
        m = np.zeros((int(keep.sum()) + 1, cont.size + 1))
        m[0, 1:], m[1:, 0], m[1:, 1:] = cont, r_axis[keep], tab[keep]
        path = os.path.join(Emap["emap_save_dir"], Emap[key])
        np.savetxt(path, m)
        paths.append(path)
    return paths


def _to_float_list(x):
    if isinstance(x, (list, tuple, np.ndarray)):
        return [float(v) for v in x]
    if isinstance(x, str):
        s = x.strip()
        return [float(v.strip()) for v in s.split(",") if v.strip()] if "," in s else [float(s)]
    return [float(x)]


def build_frequency_curve_from_config(config, Emap, machine):
    """(t, f, V0, E) on the uniform time grid; also written to
    config['filename'] with the header the field classes read."""
    save_path = config["filename"]
    e_inj, e_ext = float(machine["energy_inj"]), float(machine["energy_ext"])
    gap_model = str(Emap["gap_model"]).lower()
    if gap_model not in ("fixed", "spiral"):
        raise ValueError("Emap['gap_model'] must be 'fixed' or 'spiral'.")
    ngap, harmonic = int(Emap["Ngap"]), int(Emap["harmonic"])
    t_start = float(Emap["t_start"])
    if ngap <= 0:
        raise ValueError("Emap['Ngap'] must be a positive integer.")
    if t_start > 0.0:
        raise ValueError("Emap['t_start'] should be <= 0.")
    gap_az = _to_float_list(Emap["gap_azimuth"])
    if gap_model == "fixed" and len(gap_az) != ngap:
        raise ValueError("For fixed gap model, len(gap_azimuth) must equal Emap['Ngap'].")
    if gap_model == "spiral":
        gap_az = [360.0 * i / ngap for i in range(ngap)]
    rf_shift = [(-360.0 * harmonic * i / ngap) for i in range(ngap)]
    seo_dir = os.path.join(Emap["Bmap_fold"], Emap.get("SEO_FileName", "resultsSEO_noError"))
    ek_nodes = np.array([])
    phi_nodes_unw = phi_nodes = r_nodes = np.array([])
    alpha_deg = r_ref = s_ref_deg = np.nan
    phi_gap_of_ek = None
    if gap_model == "spiral":
        alpha_deg, r_ref, s_ref_deg = float(Emap["spiral_alpha_deg"]), float(Emap["spiral_r_ref"]), \
            float(Emap["spiral_s_ref_deg"])
        phi_gap_of_ek, ek_nodes, phi_nodes_unw, r_nodes = build_phi_gap_interpolator_from_SEO(
            os.path.join(seo_dir, "SEO_r.txt"), alpha_deg, r_ref, s_ref_deg, float(Emap["spiral_rmin_calc"]),
            float(Emap["spiral_rmax_calc"]), int(Emap["spiral_Nphi"]))
        phi_nodes = wrap_0_2pi(phi_nodes_unw)
    seo_ini = os.path.join(seo_dir, "SEO_ini.txt")
    if not os.path.exists(seo_ini):
        raise FileNotFoundError(f"{seo_ini} is missing: the revolution frequency comes from the closed orbits")
    sd = np.loadtxt(seo_ini, skiprows=2, ndmin=2)
    o = np.argsort(sd[:, 1])
    f_rev_of = _linear_extrap(sd[o, 1], sd[o, 4])

    ph0, ph1 = np.deg2rad(float(Emap["acc_phase_start"])), np.deg2rad(float(Emap["acc_phase_end"]))
    pg0, pg1 = int(Emap["phase_ramp_start"]) * ngap, int(Emap["phase_ramp_end"]) * ngap
    v0s, v0e = float(Emap["acc_voltage_start"]), float(Emap["acc_voltage_end"])
    vg0, vg1 = int(Emap["voltage_ramp_start"]) * ngap, int(Emap["voltage_ramp_end"]) * ngap

    def ramp(i, g0, g1, a, b):
        if i < g0:
            return a
        if i <= g1:
            return a + (b - a) * ((i - g0) / max(1, g1 - g0))
        return b

    E, t, V, f = [e_inj], [0.0], [ramp(0, vg0, vg1, v0s, v0e)], []
    f_inj = float(f_rev_of(e_inj))
    dt0 = (1.0 / f_inj) / ngap
    f.append(harmonic / (ngap * dt0) + (ramp(1, pg0, pg1, ph0, ph1) - ramp(0, pg0, pg1, ph0, ph1)) / (2 * np.pi * dt0))
    gap_angle = 2.0 * np.pi / ngap
    turn = gap = 0
    for ev in range(int(Emap["num_turns"]) * ngap):
        phs, v0 = ramp(ev, pg0, pg1, ph0, ph1), ramp(ev, vg0, vg1, v0s, v0e)
        e_new = E[-1] + v0 * np.sin(phs) / 1e6
        f_new = float(f_rev_of(e_new))
        dth = float(wrap_pm_pi(phi_gap_of_ek(e_new) - phi_gap_of_ek(E[-1]))) if gap_model == "spiral" else 0.0
        dt = (1.0 / f_new) / ngap * ((gap_angle + dth) / gap_angle)
        dphi = ramp(ev + 1, pg0, pg1, ph0, ph1) - phs
        E.append(e_new)
        t.append(t[-1] + dt)
        V.append(v0)
        f.append(harmonic / (ngap * dt) + dphi / (2.0 * np.pi * dt))
        gap = (gap + 1) % ngap
        if gap == 0:
            turn += 1
        if e_new > e_ext or turn >= int(Emap["num_turns"]):
            break
    t, f, V, E = (np.array(a, dtype=np.float64) for a in (t, f, V, E))
    if f.size == 0 or np.max(f) <= 0:
        raise RuntimeError("invalid frequency array")
    dtc = 0.5 / np.max(f)
    t_end = max(t[-1], t_start)
    ts = np.concatenate(([t_start], t))
    tu = np.arange(t_start, t_end + 0.5 * dtc, dtc, dtype=np.float64)
    out = np.column_stack((tu, np.interp(tu, ts, np.concatenate(([f[0]], f))),
                           np.interp(tu, ts, np.concatenate(([V[0]], V))),
                           np.interp(tu, ts, np.concatenate(([E[0]], E)))))
    header = [
        f"gap_model = {gap_model}", f"Ngap = {ngap}", f"acc_voltage_paint = {v0s}", f"acc_voltage_acc = {v0e}",
        f"voltage_ramp_start = {int(Emap['voltage_ramp_start'])}", f"voltage_ramp_end = {int(Emap['voltage_ramp_end'])}",
        f"voltage_gap_start = {vg0}", f"voltage_gap_end = {vg1}",
        f"acc_phi_paint = {float(Emap['acc_phase_start'])}", f"acc_phi_acc = {float(Emap['acc_phase_end'])}",
        f"turn_ramp_start = {int(Emap['phase_ramp_start'])}", f"turn_ramp_end = {int(Emap['phase_ramp_end'])}",
        f"phase_gap_start = {pg0}", f"phase_gap_end = {pg1}", f"gap_azimuth = {gap_az}", f"rf_shift = {rf_shift}",
        f"gap_width = {float(Emap['gap_width'])}", f"E_rmin = {float(Emap['rmin'])}",
        f"E_rmax = {float(Emap['rmax'])}", f"harmonic = {harmonic}", f"spiral_alpha_deg = {alpha_deg}",
        f"spiral_r_ref = {r_ref}", f"spiral_s_ref_deg = {s_ref_deg}", f"spiral_gap_node_count = {ek_nodes.size}",
        f"Ek_gap = {ek_nodes.tolist()}", f"r_gap = {r_nodes.tolist()}",
        f"phi_gap_deg = {np.rad2deg(phi_nodes).tolist()}",
        f"phi_gap_unwrapped_deg = {np.rad2deg(phi_nodes_unw).tolist()}", f"dt_uniform = {dtc:.6e}",
        f"Er_map_file = {Emap.get('Er_map_file', 'Er_coef.txt')}",
        f"Ez_map_file = {Emap.get('Ez_map_file', 'Ez_coef.txt')}",
        f"Ephi_map_file = {Emap.get('Ephi_map_file', 'Ephi_coef.txt')}",
        "#Time (s)        Frequency (Hz)    V0 (V)    Energy (MeV)",
    ]
    d = os.path.dirname(save_path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(save_path, "w", encoding="utf-8") as fh:
        fh.write("\n".join(header) + "\n")
        np.savetxt(fh, out, fmt=config.get("save_precision", "%.8e"), delimiter="\t")
    return out


def main_Emap(json_path):
    """The main_Emap flow on a config_Emap.json (no plots)."""
    with open(json_path, encoding="utf-8") as f:
        cfg = json.load(f)
    config, emap, machine = cfg["config"], cfg["Emap"], cfg["machine"]
    if str(emap["gap_model"]).lower() not in ("fixed", "spiral"):
        raise ValueError("Emap['gap_model'] must be 'fixed' or 'spiral'.")
    out = {}
    if config.get("build_frequency_curve", True):
        out["frequency_curve"] = build_frequency_curve_from_config(config, emap, machine)
    if config.get("build_spiral_map", False) and str(emap["gap_model"]).lower() == "spiral":
        out["maps"] = build_spiral_map_from_config(config, emap)
    return out
