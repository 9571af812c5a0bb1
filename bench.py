"""Throughput benchmark of the FFAG tracking hot path on one MI355X.

Workload (headline): 12-sector scaling FFAG (r0 = 8.5 m, k = 7.6, 300 MeV
protons, order-2 Taylor maps on an r 8.0-9.3 m x one-sector grid), 8M
macro-particles (seeded Gaussian bunch, uniform azimuth, 1e13 protons), 2.5D
space charge on a (ns, nx, nz) = (512, 256, 256) grid at +-4 sigma every 20
push steps, two RF gaps 180 degrees apart at 100 kV with a linear frequency
ramp, the aperture r 8.0-9.3 m |z| < 5 cm, Boris pusher, dt = 0.05 ns.

One bench step = 1000 Boris push steps: the reference's progress / stop-check
interval (src/FFAG_track.py:518-577).  It holds 50 SC solves, one stop-check
reduction and, in the tiers that order particles, five azimuth sorts (every
200 steps, the FP32 tier's default; the accuracy and tune legs sort at the
same interval unless --accuracy-sort / --tune-sort says otherwise).

The headline value is the FP32-field tier (FP64 state, fused steps; each
SC solve stops after the field and the push launch that follows applies its
gather and endpoint roll).  Beside
it, on the same workload: the FP64 map-gather path in the reference's
particle order (the parity path), the FP64 path with the same azimuth
ordering as the FP32 tier, and that ordered FP64 path with its steps fused
into one launch per block (bit-identical to it), so the speed-ups separate
what FP32 buys from what ordering and fusion buy; the FP32 tier's accuracy
against FP64 (centroid, rms emittance) and its field error against the FP64
field at sampled points.

Further legs (sub-objects of the JSON line):
* ``surrogate`` -- the field network trained on the GPU on 4M samples of the
  order-4 Taylor-map field, driving the FP32 tier on the same 8M workload
  with order-4 maps as fallback; beside it the FP32 and FP64 order-4 map
  tiers; training time, field error, matrix-core throughput of the network;
* ``map_generator`` -- Taylor-map generation on the full map grid;
* ``spiral_cycle`` (``--cycle-steps K``) -- the 32M-particle spiral-sector
  acceleration cycle (order-6 maps, spiral gaps, (1024, 512, 512) grid),
  FP32 tier against the fused, ordered FP64 path, timed after injection;
* ``cpu_baseline`` -- the OpenMP FP64 oracle on the host cores, on a sample
  of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_PART = 8_000_000
GRID = (512, 256, 256, 4.0)      # (ns, nx, nz, sigma_ratio)
SC_STEP = 20
PUSHES = 1000                    # push steps per bench step
SORT_EVERY = 200                 # steps between azimuth sorts in the ordered tiers (the FP32 tier's default)
DT = 0.05e-9
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
# dense matrix-core FLOPs of one network evaluation: 2 * (64*128 + 3*128*128 + 128*3)
MLP_FLOP = 2 * (64 * 128 + 3 * 128 * 128 + 128 * 3)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_fused_push_8M.json")
PUSH_KERNEL = "k_boris_fused<float, false, 1, 0>"  # the FP32 push that applies the gather from the cubes
VALU_PROFILE = os.path.join(ROOT, "profiles", "valu_fused_push_8M.json")


def measured_traffic(kernel):
    """HBM bytes per launch (read + written) of a kernel of this workload,
    from the two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE, scaled by a
    256 MiB copy) summarised in profiles/; None when no summary is present."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            k = json.load(f)["kernels"][kernel]
        return k["read_bytes"] + k["write_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def measured_valu(kernel):
    """VALUBusy / VALUUtilization (rocprofv3 derived counters) of a kernel of
    this workload, from the summary in profiles/; None when absent."""
    try:
        with open(VALU_PROFILE) as f:
            k = json.load(f)["kernels"][kernel]
        return {"valu_busy_pct": k["VALUBusy"], "valu_utilization_pct": k["VALUUtilization"]}
    except (OSError, KeyError, ValueError):
        return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def machines(max_order, tmpdir):
    """FFAG_GlobalParameters of the workload with Taylor maps to max_order."""
    from pyffa_amd import workloads as W
    return W.global_parameters("stable", max_order=max_order, rf=dict(v0=100e3), aperture=W.APERTURE, tmpdir=tmpdir)


def snapshot(B):
    return {k: B.ordered(k) for k in ("r", "vr", "z", "vz", "fidot", "flags")}


def moments(S, alive):
    """Centroid and rms emittances (r, vr/v) and (z, vz/v) over the mask."""
    import torch
    r, z = S["r"][alive], S["z"][alive]
    v = torch.sqrt(S["vr"][alive] ** 2 + S["vz"][alive] ** 2 + (S["r"][alive] * S["fidot"][alive]) ** 2)
    xp, zp = S["vr"][alive] / v, S["vz"][alive] / v

    def emit(a, b):
        a = a - a.mean()
        b = b - b.mean()
        return float(torch.sqrt((a * a).mean() * (b * b).mean() - (a * b).mean() ** 2))
    return dict(r_mean=float(r.mean()), z_mean=float(z.mean()), emit_r=emit(r, xp), emit_z=emit(z, zp),
                alive=int(alive.sum()))


def accuracy(Sa, Sb, steps):
    """Moments of tier a against tier b over the particles alive in both."""
    both = ((Sa["flags"] & 2) != 0) & ((Sb["flags"] & 2) != 0)
    ma, mb = moments(Sa, both), moments(Sb, both)
    return {"centroid_r_abs_m": abs(ma["r_mean"] - mb["r_mean"]), "centroid_z_abs_m": abs(ma["z_mean"] - mb["z_mean"]),
            "emit_r_rel": abs(ma["emit_r"] / mb["emit_r"] - 1.0), "emit_z_rel": abs(ma["emit_z"] / mb["emit_z"] - 1.0),
            "alive_a": int(((Sa["flags"] & 2) != 0).sum()), "alive_b": int(((Sb["flags"] & 2) != 0).sum()),
            "compared": ma["alive"], "after_steps": steps}


def run_tier(precision, M, gp, warmup, steps, grid=GRID, surrogate=None, sort_every=None, marcosize=None,
             fuse_fp64=False):
    """Tracks the bunch ``warmup`` untimed then ``steps`` timed bench steps
    (PUSHES push steps each) through the tracking loop; returns the bunch,
    the tracker, the timed seconds, the final state in input order and the
    average kernel time per push launch and per SC solve inside the timed
    run (HIP events on the launch stream)."""
    import torch
    from pyffa_amd import workloads as W
    from pyffa_amd.FFAG_Bunch import FFAG_Bunch
    from pyffa_amd.FFAG_track import FFAG_RungeKutta
    B = FFAG_Bunch(M, marcosize=marcosize or 1e13 / M.shape[0], sc_grid_size=grid, BunchType="Boris",
                   precision=precision)
    W.set_bunch_seo(B)
    rk = FFAG_RungeKutta()

    def lp(first, last, timer=None):
        d = dict(stop_condition=dict(max_stepsN=last * PUSHES, max_turn=1e9, max_time=1.0),
                 time_step=DT, step_dumps=None, enable_SC=True, SC_type=1, SC_step=SC_STEP,
                 precision=precision, verbose=False, start_step=first * PUSHES)
        if surrogate is not None:
            d["surrogate"] = surrogate
        if sort_every is not None:
            d["sort_every"] = sort_every
        if fuse_fp64:
            d["fuse_fp64"] = True
        if timer is not None:
            d["kernel_timer"] = timer
        return d
    t = 0.0
    if warmup:
        rk.rk4_solve_dt_bunch3_withSC_boris(None, t, B, gp, lp(0, warmup))
        t = rk.last_run["t"]
    timer = LaunchTimer(torch.cuda.current_stream())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rk.rk4_solve_dt_bunch3_withSC_boris(None, t, B, gp, lp(warmup, warmup + steps, timer))
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return B, rk, el, snapshot(B), timer.result()


class LaunchTimer:
    """HIP events on the launch stream around every push call and SC solve
    of a tracking run (the loop's ``kernel_timer`` hook): average duration
    per kernel launch over the timed region."""

    def __init__(self, stream):
        self.stream = stream
        self.open = {}
        self.pairs = {"push": [], "sc": []}

    def begin(self, name):
        import torch
        ev = torch.cuda.Event(enable_timing=True)
        ev.record(self.stream)
        self.open[name] = ev

    def end(self, name, launches):
        import torch
        ev = torch.cuda.Event(enable_timing=True)
        ev.record(self.stream)
        self.pairs[name].append((self.open.pop(name), ev, launches))

    def result(self):
        """{name: average ms per kernel launch} (after the stream is synchronised)."""
        out = {}
        for name, pr in self.pairs.items():
            if pr:
                out[name] = sum(a.elapsed_time(b) for a, b, _ in pr) / sum(k for _, _, k in pr)
        out["push_launches"] = sum(k for _, _, k in self.pairs["push"])
        out["sc_solves"] = len(self.pairs["sc"])
        return out


def cpu_baseline(M, gp, n_sample, blocks, n_total):
    """The FP64 OpenMP oracle on the host cores, on a bounded sample of the
    workload: its first ``n_sample`` particles (macro-particle charge of the
    full bunch) for ``blocks`` x SC_STEP Boris steps with one SC solve per
    SC_STEP steps on the workload's grid."""
    from oracle import adapters as A
    from oracle import oracle as O
    from pyffa_amd import workloads as W
    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    O.lib().ora_set_threads(threads)
    S = M[:n_sample].copy()
    tr = O.Tracker(A.machine_from(gp), S, enable_sc=True, sc_step=SC_STEP, grid=GRID, marcosize=1e13 / n_total,
                   seo=A.seo_from(W.seo("stable")))
    t0 = time.perf_counter()
    tr.run(0.0, DT, blocks * SC_STEP)
    el = time.perf_counter() - t0
    return dict(value=n_sample * blocks * SC_STEP / el, unit="particle-steps/s", cores=threads, kind="port",
                sample=f"first {n_sample} of the {n_total} particles x {blocks * SC_STEP} Boris steps with "
                       f"{blocks} SC solves on the {GRID[0]}x{GRID[1]}x{GRID[2]} grid (FP64 OpenMP oracle), "
                       f"{el:.1f} s")


def surrogate_leg(M, tmp, warmup, steps, train_iters, zt=0.01):
    """Train the network on the GPU, then run the FP32 tier with it against
    the FP32 and FP64 order-4 map tiers (both azimuth-ordered) on the
    headline workload."""
    import torch
    from pyffa_amd.FFAG_Surrogate import FFAG_FieldSurrogate
    gp4 = machines(4, tmp)
    sur = FFAG_FieldSurrogate(gp4, zt=zt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sur.build_dataset(4_000_000, seed=1)
    hist = sur.train(iters=train_iters, batch=8192, lr=2e-3, seed=2, log_every=max(1, train_iters // 4))
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    err = sur.field_error(n=1_000_000, seed=9)
    log(f"surrogate trained in {t_train:.1f} s, field error {err}")
    # matrix-core throughput of the network alone (4M points, HIP events)
    pts = torch.stack(sur.sample_points(1 << 22, 3)).float().contiguous()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    sur.eval(pts)
    ev[0].record()
    for _ in range(10):
        sur.eval(pts)
    ev[1].record()
    ev[1].synchronize()
    t_eval = ev[0].elapsed_time(ev[1]) / 10 * 1e-3
    tflops = pts.shape[1] * MLP_FLOP / t_eval / 1e12
    out = {"train_seconds": t_train, "train_samples": 4_000_000, "train_iters": train_iters, "batch": 8192,
           "loss_history": hist, "field_error": err, "zt_m": sur.zt,
           "mfma": {"kernel": "k_mlp_eval (4M points)", "ms": t_eval * 1e3, "achieved_tflops": tflops,
                    "peak_tflops": BF16_PEAK_TFLOPS, "frac": tflops / BF16_PEAK_TFLOPS}}
    ps = M.shape[0] * steps * PUSHES
    B, rk, el_sur, S_sur, k_sur = run_tier("fp32", M, gp4, warmup, steps, surrogate=sur)
    del B, rk
    B, rk, el_map32, S_map32, _ = run_tier("fp32", M, gp4, warmup, steps)
    del B, rk
    B, rk, el_map64, S_map64, _ = run_tier("fp64", M, gp4, warmup, steps, sort_every=SORT_EVERY)
    del B, rk
    after = (warmup + steps) * PUSHES
    out.update({"value": ps / el_sur, "kernel_ms": k_sur,
                "map_order4_fp32_value": ps / el_map32, "map_order4_fp64_value": ps / el_map64,
                "speedup_vs_fp64_order4": el_map64 / el_sur,
                "accuracy_vs_fp64_order4": accuracy(S_sur, S_map64, after),
                "fp32_map_accuracy_vs_fp64_order4": accuracy(S_map32, S_map64, after)})
    log(f"surrogate tier {ps / el_sur:.3e}, fp32 order-4 {ps / el_map32:.3e}, fp64 order-4 {ps / el_map64:.3e}")
    return out


def orbit_moments(S, seo):
    """Centroid and rms emittances about the closed orbit: x = r - r0(Ek, phi),
    x' = vr / v_phi - pr0(Ek, phi) (the SEO tables the SC solve uses; pr0 is
    the orbit slope dr/(r dphi)), and (z, vz / v_phi), over the alive
    particles."""
    import numpy as np
    fl = S["flags"].cpu().numpy()
    al = (fl & 2) != 0
    r, z = S["r"].cpu().numpy()[al], S["z"].cpu().numpy()[al]
    vr, vz = S["vr"].cpu().numpy()[al], S["vz"].cpu().numpy()[al]
    vf = r * S["fidot"].cpu().numpy()[al]
    fi = np.mod(S["fi"].cpu().numpy()[al], 2 * np.pi)
    g = 1.0 / np.sqrt(1.0 - (vr * vr + vz * vz + vf * vf) / 2.99792458e8 ** 2)
    ek = (g - 1.0) * 938.2723

    def bil(tab):
        fx = np.clip((ek - seo.Ek_axis[0]) / (seo.Ek_axis[1] - seo.Ek_axis[0]), 0, seo.Ek_axis.size - 1.000001)
        fy = np.clip((fi - seo.fi_axis[0]) / (seo.fi_axis[1] - seo.fi_axis[0]), 0, seo.fi_axis.size - 1.000001)
        i, j = np.floor(fx).astype(int), np.floor(fy).astype(int)
        tx, ty = fx - i, fy - j
        return ((1 - tx) * (1 - ty) * tab[i, j] + (1 - tx) * ty * tab[i, j + 1] + tx * (1 - ty) * tab[i + 1, j]
                + tx * ty * tab[i + 1, j + 1])
    x = r - bil(seo.r_matrix)
    xp = vr / vf - bil(seo.pr_matrix)
    zp = vz / vf

    def emit(a, b):
        a, b = a - a.mean(), b - b.mean()
        return float(np.sqrt(np.mean(a * a) * np.mean(b * b) - np.mean(a * b) ** 2))
    return dict(x_mean=float(x.mean()), z_mean=float(z.mean()), emit_x=emit(x, xp), emit_z=emit(z, zp),
                alive=int(al.sum()))


def _grid_arg(text):
    """'ns,nx,nz' -> (ns, nx, nz, sigma_ratio) like GRID, or GRID itself."""
    return GRID if not text else tuple(int(v) for v in text.split(",")) + (GRID[3],)


def accuracy_leg(tmp, n, turns, tiers, chunk_turns=50, grid=GRID, sort_every=SORT_EVERY):
    """Tracks the same bunch for ``turns`` turns (SC, RF, aperture) in the
    requested runs: "fp32" (FP32 tier), "fp64" (FP64 path) and "fp64p" (FP64
    path from r and z moved by one ulp: the problem's own sensitivity).  The
    closed-orbit moments are recorded every ``chunk_turns`` turns; the
    relative emittance differences of fp32 and fp64p against fp64 are
    reported side by side.  Every run reorders its particles by azimuth
    every ``sort_every`` steps (for the FP64 path this changes only the partition of
    the SC moment sums, a rounding-level difference like the one fp64p
    measures; ``sort_every`` 1000 reproduces the runs recorded before the
    FP32 tier's default became 200).  The FP64 runs fuse each block of steps
    into one launch (bit-identical to one launch per step)."""
    import numpy as np
    from pyffa_amd.FFAG_Bunch import FFAG_Bunch
    from pyffa_amd.FFAG_track import FFAG_RungeKutta
    from pyffa_amd import workloads as W
    gp = machines(2, tmp)
    seo = W.seo("stable")
    M0 = W.bunch("stable", n=n, seed=3)
    steps_per_turn = seo.T0[3] / DT
    total = int(round(turns * steps_per_turn / SC_STEP))
    n_chunks = max(1, int(round(turns / chunk_turns)))
    out = {"particles": n, "turns": turns, "sc_grid": list(grid[:3]), "steps": total * SC_STEP, "sort_every": sort_every,
           "moments": "about the closed orbit r0(Ek, phi), pr0(Ek, phi) of the SEO tables"}
    for run in tiers:
        prec = "fp32" if run == "fp32" else "fp64"
        M = M0.copy()
        if run == "fp64p":
            M[:, 0] = np.nextafter(M[:, 0], np.inf)
            M[:, 2] = np.nextafter(M[:, 2], np.inf)
        B = FFAG_Bunch(M, marcosize=1e13 / n, sc_grid_size=grid, BunchType="Boris", precision=prec)
        W.set_bunch_seo(B)
        rk = FFAG_RungeKutta()
        t, done, hist = 0.0, 0, []
        t0 = time.perf_counter()
        for c in range(n_chunks):
            last = total * (c + 1) // n_chunks
            lp = dict(stop_condition=dict(max_stepsN=last * SC_STEP, max_turn=1e9, max_time=1.0), time_step=DT,
                      step_dumps=None, enable_SC=True, SC_type=1, SC_step=SC_STEP, precision=prec, verbose=False,
                      start_step=done * SC_STEP, sort_every=sort_every, fuse_fp64=prec == "fp64")
            rk.rk4_solve_dt_bunch3_withSC_boris(None, t, B, gp, lp)
            t, done = rk.last_run["t"], last
            S = {k: B.ordered(k) for k in ("r", "vr", "z", "vz", "fi", "fidot", "flags")}
            m = orbit_moments(S, seo)
            hist.append(dict(turns=done * SC_STEP / steps_per_turn, **m))
            log(f"accuracy {run}: {hist[-1]['turns']:.0f} turns, alive {m['alive']}, emit x {m['emit_x']:.6e} "
                f"z {m['emit_z']:.6e}, {time.perf_counter() - t0:.0f} s")
        out[run] = {"history": hist, "seconds": time.perf_counter() - t0}
        del B, rk
    if "fp64" in out:
        ref = out["fp64"]["history"]
        for run in ("fp32", "fp64p"):
            if run in out:
                d = [dict(turns=a["turns"], emit_x_rel=a["emit_x"] / b["emit_x"] - 1.0,
                          emit_z_rel=a["emit_z"] / b["emit_z"] - 1.0, alive=a["alive"] - b["alive"])
                     for a, b in zip(out[run]["history"], ref)]
                out[f"{run}_vs_fp64"] = {"per_chunk": d,
                                         "max_abs_emit_x_rel": max(abs(e["emit_x_rel"]) for e in d),
                                         "max_abs_emit_z_rel": max(abs(e["emit_z_rel"]) for e in d),
                                         "final": d[-1]}
    return out


def cycle_leg(tmp, n, steps, precisions=("fp32", "fp64")):
    """The spiral-sector acceleration cycle (35 degree spiral, order-6 GPU
    maps, two 200 kV spiral gaps, painted injection of 4 sub-bunches, the
    (1024, 512, 512) SC grid): the job runs through ``main_track.main`` until
    the whole bunch is injected (200 ns after injection starts), then
    ``steps`` more steps are timed, the FP32 tier and the FP64 path with its
    steps fused and the same azimuth ordering."""
    import contextlib
    import torch
    from pyffa_amd import FFAG_Distribution as FD
    from pyffa_amd import main_track, workloads as W
    from pyffa_amd.FFAG_track import FFAG_RungeKutta
    inj_step = 4000                      # tracking starts 200 ns before injection
    out = {"particles": n, "sc_grid": [1024, 512, 512], "max_order": 6, "timed_steps": steps}
    for prec in precisions:
        folder = os.path.join(tmp, f"cycle_{prec}")
        os.makedirs(folder, exist_ok=True)
        path, _ = W.spiral_cycle_job(folder, n_particles=n, grid=(1024, 512, 512, 4.0), max_steps=inj_step + 4200,
                                     precision=prec, paint_bunches=4, paint_interval_ns=50.0, bunch_length_ns=50.0)
        FD.set_seed(3)
        t0 = time.perf_counter()
        with contextlib.redirect_stdout(sys.stderr):          # the job's progress lines
            rk, B, gp = main_track.main(path)
        torch.cuda.synchronize()
        log(f"cycle {prec}: job set up and {inj_step + 4200} steps in {time.perf_counter() - t0:.1f} s, "
            f"injected {B.InjectedNum_Local}")
        s0, t = rk.last_run["steps"], rk.last_run["t"]
        timer = LaunchTimer(torch.cuda.current_stream())
        lp = dict(stop_condition=dict(max_stepsN=s0 + steps, max_turn=1e9, max_time=1.0), time_step=DT,
                  step_dumps=None, enable_SC=True, SC_type=1, SC_step=SC_STEP, precision=prec, verbose=False,
                  start_step=s0, sort_every=SORT_EVERY, fuse_fp64=prec == "fp64", kernel_timer=timer)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        FFAG_RungeKutta().rk4_solve_dt_bunch3_withSC_boris(None, t, B, gp, lp)
        torch.cuda.synchronize()
        el = time.perf_counter() - t1
        out[prec] = {"particle_steps_per_s": n * steps / el, "seconds": el,
                     "alive": int(((B.flags & 2) != 0).sum()), "kernel_ms": timer.result()}
        log(f"cycle {prec}: {out[prec]}")
        del rk, B, gp
    if "fp32" in out and "fp64" in out:
        out["speedup_fp32_vs_fp64_fused_sorted"] = (out["fp32"]["particle_steps_per_s"] /
                                                   out["fp64"]["particle_steps_per_s"])
    return out


def mapgen_leg(orders=(2, 6), reps=5, cpu_rows=2, cpu_order=4, cpu=True):
    """Taylor-map generation on the full map grid of the 12-sector lattice
    (r 8.0-9.3 m at 1 mm: 1301 radii; 30 deg at 0.005 deg: 6000 azimuths)
    with a 35 deg spiral: the FP64 stacks of every order up to max_order
    (3 (max_order+1) x 1301 x 6000 doubles) and the FP32 tracking table,
    timed with HIP events on the generator's stream.  The fill is
    write-bound: its algorithmic bytes are the map bytes it stores.  The
    CPU figure is the oracle's restatement of the reference generator on
    ``cpu_rows`` radii at order ``cpu_order``, scaled to the grid."""
    import numpy as np
    import torch
    from numpy.polynomial import Polynomial
    from pyffa_amd import FFAG_HighOrder as H
    from pyffa_amd import workloads as W
    cfg = W.lattice("stable").to_config()["Bmap"]
    r, fi = H.grid_from_config(cfg)
    s_fit = Polynomial.fit(r, H.calculate_s(r, 0.0, 35.0), 5)
    conv = H.BmapParams_to_ConvexParams(cfg)
    coef = tuple(cfg["coefficients_tupel"])
    out = {"grid": [int(r.size), int(fi.size)], "spiral_deg": 35.0}
    for mo in orders:
        t0 = time.perf_counter()
        gen = H.TaylorMapGenerator(mo, r, fi, conv, s_fit, coef)
        host_s = time.perf_counter() - t0
        res = {"host_weights_s": host_s}
        for layout, dt in (("stacks", torch.float64), ("tracking", torch.float32)):
            buf = torch.empty(gen.output_shape(layout), dtype=dt, device="cuda")
            gen.run(layout, dt, out=buf)
            stream = torch.cuda.current_stream()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record(stream)
            for _ in range(reps):
                gen.run(layout, dt, out=buf)
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            nbytes = buf.numel() * buf.element_size()
            res[layout] = {"dtype": str(dt).replace("torch.", ""), "bytes": nbytes, "ms": ms,
                           "write_GBps": nbytes / ms / 1e6, "frac_of_8TBps": nbytes / ms / 1e6 / 8000.0}
            del buf
        out[f"order_{mo}"] = res
        log(f"map generator order {mo}: stacks {res['stacks']['ms']:.3f} ms "
            f"({res['stacks']['write_GBps']:.0f} GB/s), tracking f32 {res['tracking']['ms']:.3f} ms")
    if not cpu:
        return out
    from oracle import taylor_gen as T
    poles = [(x0, x1, w0, sg * m) for x0, x1, w0, (m, sg) in conv]
    t0 = time.perf_counter()
    T.generate_maps(cpu_order, r[:cpu_rows], fi, poles, s_fit, coef)
    per_row = (time.perf_counter() - t0) / cpu_rows
    out["cpu_baseline"] = {"kind": "port", "cores": 1, "order": cpu_order, "sample_rows": cpu_rows,
                           "seconds_per_row": per_row, "seconds_full_grid": per_row * r.size}
    log(f"map generator oracle: {per_row:.2f} s per radius at order {cpu_order}")
    return out


def tune_leg(tmp, n, turns, n_track, runs=("bare", "fp64", "fp32", "fp64p"), progress=False, grid=GRID,
             save_dir=None, load_dir=None, sort_every=SORT_EVERY):
    """Incoherent tunes with space charge: the bunch of ``accuracy_leg`` is
    tracked ``turns`` + 1.5 turns in each run ("bare": FP64 without SC;
    "fp64", "fp32": with SC; "fp64p": FP64 with SC from r and z moved by one
    ulp, the problem's own sensitivity); a PositionDump at azimuth 0 records
    the first ``n_track`` particles every turn and the FFT tune estimator
    gives their (qx, qz) about the closed orbit.  The tune shift of a
    particle is its tune with SC minus its bare tune; the shifts of "fp32"
    and "fp64p" are compared with those of "fp64" particle by particle.
    ``save_dir``: each run's per-particle tunes are written there as
    ``tune_<run>.json``; ``load_dir``: runs saved there by an earlier call
    (same workload) join the comparison, so the runs can be split over
    several calls.  ``sort_every``: steps between azimuth sorts in every run
    (1000 reproduces the runs recorded before the FP32 tier's default became
    200); the FP64 runs fuse each block of steps into one launch
    (bit-identical to one launch per step)."""
    import numpy as np
    from pyffa_amd import diagnostics as D
    from pyffa_amd.FFAG_Bunch import FFAG_Bunch
    from pyffa_amd.FFAG_dump import Dumps, PositionDump
    from pyffa_amd.FFAG_track import FFAG_RungeKutta
    from pyffa_amd import workloads as W
    gp = machines(2, tmp)
    seo = W.seo("stable")
    M0 = W.bunch("stable", n=n, seed=3)
    steps_per_turn = seo.T0[3] / DT
    total = int(round((turns + 1.5) * steps_per_turn / SC_STEP)) * SC_STEP
    out = {"particles": n, "turns": turns, "tracked": n_track, "sc_grid": list(grid[:3]), "steps": total,
           "estimator": "FFT of x + i x'/scale, 8x zero padding, parabolic peak; x about r0(Ek, phi)"}
    q = {}
    if load_dir:
        for run in ("bare", "fp64", "fp32", "fp64p"):
            path = os.path.join(load_dir, f"tune_{run}.json")
            if run in runs or not os.path.exists(path):
                continue
            with open(path) as f:
                d = json.load(f)
            qx, qz = np.asarray(d["qx"]), np.asarray(d["qz"])
            q[run] = dict(zip(d["ids"], zip(d["qx"], d["qz"])))
            out[run] = {"seconds": d["seconds"], "particles_analysed": len(d["ids"]), "loaded_from": path,
                        "qx_mean": float(qx.mean()), "qz_mean": float(qz.mean()),
                        "qx_rms": float(qx.std()), "qz_rms": float(qz.std())}
    for run in runs:
        prec = "fp32" if run == "fp32" else "fp64"
        M = M0.copy()
        if run == "fp64p":
            M[:, 0] = np.nextafter(M[:, 0], np.inf)
            M[:, 2] = np.nextafter(M[:, 2], np.inf)
        B = FFAG_Bunch(M, marcosize=1e13 / n, sc_grid_size=grid, BunchType="Boris", precision=prec)
        W.set_bunch_seo(B)
        folder = os.path.join(tmp, f"tune_{run}")
        pd = PositionDump(0.0, -1.0, 1.0, np.arange(n_track), save_folder=folder, max_records=turns + 8)
        dumps = Dumps()
        dumps.add_dump(pd)
        rk = FFAG_RungeKutta()
        t0 = time.perf_counter()
        log(f"tune {run}: tracking {n} particles for {total} steps")
        lp = dict(stop_condition=dict(max_stepsN=total, max_turn=1e9, max_time=1.0), time_step=DT,
                  step_dumps=dumps, enable_SC=run != "bare", SC_type=1, SC_step=SC_STEP, precision=prec,
                  verbose=progress, sort_every=sort_every, fuse_fp64=prec == "fp64")
        rk.rk4_solve_dt_bunch3_withSC_boris(None, 0.0, B, gp, lp)
        pd.save_all_particles()
        ids, qx, qz = D.particle_tunes(D.load_position_dump(folder, 0.0), seo, turns=turns)
        q[run] = dict(zip(ids.tolist(), zip(qx.tolist(), qz.tolist())))
        out[run] = {"seconds": time.perf_counter() - t0, "particles_analysed": int(ids.size),
                    "qx_mean": float(qx.mean()), "qz_mean": float(qz.mean()),
                    "qx_rms": float(qx.std()), "qz_rms": float(qz.std())}
        log(f"tune {run}: {ids.size} particles, qx {qx.mean():.5f} qz {qz.mean():.5f}, "
            f"{out[run]['seconds']:.0f} s")
        if save_dir:
            os.makedirs(save_dir, exist_ok=True)
            with open(os.path.join(save_dir, f"tune_{run}.json"), "w") as f:
                json.dump({"ids": ids.tolist(), "qx": qx.tolist(), "qz": qz.tolist(),
                           "seconds": out[run]["seconds"], "particles": n, "turns": turns}, f)
        del B, rk
    if "bare" in q:
        shifts = {}
        for run in q:
            if run == "bare":
                continue
            common = sorted(set(q[run]) & set(q["bare"]))
            d = np.array([[q[run][g][0] - q["bare"][g][0], q[run][g][1] - q["bare"][g][1]] for g in common])
            shifts[run] = dict(zip(common, d.tolist()))
            out[run].update({"dqx_mean": float(d[:, 0].mean()), "dqz_mean": float(d[:, 1].mean()),
                             "dqx_min": float(d[:, 0].min()), "dqz_min": float(d[:, 1].min())})
        for run in ("fp32", "fp64p"):
            if run not in shifts or "fp64" not in shifts:
                continue
            common = sorted(set(shifts[run]) & set(shifts["fp64"]))
            a = np.array([shifts[run][g] for g in common])
            b = np.array([shifts["fp64"][g] for g in common])
            e = np.abs(a - b)
            out[f"{run}_vs_fp64"] = {
                "particles": len(common),
                "dq_abs_median": [float(np.median(e[:, 0])), float(np.median(e[:, 1]))],
                "dq_abs_p90": [float(np.quantile(e[:, 0], 0.9)), float(np.quantile(e[:, 1], 0.9))],
                "dq_abs_max": [float(e[:, 0].max()), float(e[:, 1].max())],
                "mean_shift_rel": [float(abs(a[:, 0].mean() / b[:, 0].mean() - 1)),
                                   float(abs(a[:, 1].mean() / b[:, 1].mean() - 1))],
                "median_shift_rel": [float(abs(np.median(a[:, 0]) / np.median(b[:, 0]) - 1)),
                                     float(abs(np.median(a[:, 1]) / np.median(b[:, 1]) - 1))]}
    return out


def field_error_leg(gp, n=1_000_000, seed=11):
    """Relative error of the FP32 tier's field (its own FP32 evaluation on
    the FP32 map) against the FP64 path's field at n points uniform over the
    map (r 8.0-9.3 m, all azimuths, |z| <= 5 cm)."""
    import numpy as np
    from pyffa_amd.FFAG_Field import field_at_points
    rng = np.random.default_rng(seed)
    r = rng.uniform(8.0, 9.3, n)
    fi = rng.uniform(0.0, 2 * np.pi, n)
    z = rng.uniform(-0.05, 0.05, n)
    b64 = field_at_points(gp, r, fi, z, precision="fp64")
    b32 = field_at_points(gp, r, fi, z, precision="fp32")
    rms = float(np.sqrt((b64 ** 2).sum(axis=0).mean()))
    rel = np.sqrt(((b32 - b64) ** 2).sum(axis=0)) / rms
    w = int(np.argmax(rel))
    return {"points": n, "region": "r 8.0-9.3 m, all azimuths, |z| <= 5 cm", "rms_B": rms,
            "rel_rms": float(np.sqrt(np.mean(rel ** 2))), "rel_p99": float(np.quantile(rel, 0.99)),
            "rel_max": float(rel.max()),
            "worst_point": {"r": float(r[w]), "fi": float(fi[w]), "z": float(z[w]),
                            "B_fp64": [float(v) for v in b64[:, w]], "B_fp32": [float(v) for v in b32[:, w]]},
            "measure": "|B_fp32 - B_fp64| / rms |B_fp64| per point (Bz, Bphi, Br), as the surrogate's field error"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help=f"timed bench steps ({PUSHES} pushes each)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--particles", type=int, default=N_PART)
    ap.add_argument("--cpu-particles", type=int, default=1_000_000)
    ap.add_argument("--cpu-blocks", type=int, default=6, help="SC blocks of 20 steps in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=None, help="timed bench steps of the FP64 legs (default --steps)")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--surrogate-steps", type=int, default=2)
    ap.add_argument("--only-surrogate", action="store_true", help="run the surrogate leg alone")
    ap.add_argument("--surrogate-zt", type=float, default=0.01, help="|z| range [m] the network is trained on")
    ap.add_argument("--train-iters", type=int, default=12000)
    ap.add_argument("--accuracy-turns", type=float, default=0.0,
                    help="also run both tiers this many turns and report the moments' deviation (off by default)")
    ap.add_argument("--accuracy-particles", type=int, default=N_PART)
    ap.add_argument("--accuracy-tiers", default="fp32,fp64,fp64p")
    ap.add_argument("--accuracy-grid", default=None, help="SC grid ns,nx,nz of the accuracy leg")
    ap.add_argument("--accuracy-chunk", type=float, default=50.0, help="turns between moment samples")
    ap.add_argument("--accuracy-sort", type=int, default=SORT_EVERY, help="steps between azimuth sorts, every run")
    ap.add_argument("--only-accuracy", action="store_true", help="run the accuracy leg alone")
    ap.add_argument("--cycle-steps", type=int, default=0,
                    help="timed steps of the 32M spiral acceleration-cycle leg (0: leg off)")
    ap.add_argument("--cycle-particles", type=int, default=32_000_000)
    ap.add_argument("--only-cycle", action="store_true", help="run the spiral-cycle leg alone")
    ap.add_argument("--no-mapgen", action="store_true", help="skip the Taylor-map generator leg")
    ap.add_argument("--only-mapgen", action="store_true", help="run the Taylor-map generator leg alone")
    ap.add_argument("--tune-turns", type=int, default=0,
                    help="also measure incoherent tune shifts over this many turns (off by default)")
    ap.add_argument("--tune-particles", type=int, default=N_PART)
    ap.add_argument("--tune-tracked", type=int, default=2000)
    ap.add_argument("--tune-grid", default=None, help="SC grid ns,nx,nz of the tune leg")
    ap.add_argument("--only-tune", action="store_true", help="run the tune leg alone")
    ap.add_argument("--tune-sort", type=int, default=SORT_EVERY, help="steps between azimuth sorts, every run")
    ap.add_argument("--tune-runs", default="bare,fp64,fp32,fp64p", help="runs of the tune leg in this call")
    ap.add_argument("--tune-save", default=None, help="directory for each run's per-particle tunes")
    ap.add_argument("--tune-load", default=None, help="directory of runs saved by an earlier call")
    ap.add_argument("--tune-progress", action="store_true",
                    help="the tune leg's tracking prints its progress lines (on stdout, before the JSON line)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 256 MiB device copy (reference for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import tempfile
    import torch
    if args.copy_calibration:
        a = torch.empty(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log("copy calibration: 256 MiB read + 256 MiB written")
    tmp = tempfile.mkdtemp()
    if args.only_mapgen:
        print(json.dumps({"map_generator": mapgen_leg(cpu=not args.no_cpu)}), flush=True)
        return
    if args.only_tune:
        print(json.dumps({"tune_shift": tune_leg(tmp, args.tune_particles, args.tune_turns, args.tune_tracked,
                                                 runs=tuple(args.tune_runs.split(",")), progress=args.tune_progress,
                                                 grid=_grid_arg(args.tune_grid), save_dir=args.tune_save,
                                                 load_dir=args.tune_load, sort_every=args.tune_sort)}), flush=True)
        return
    if args.only_cycle:
        print(json.dumps({"spiral_cycle": cycle_leg(tmp, args.cycle_particles, args.cycle_steps or 200)}),
              flush=True)
        return
    if args.only_accuracy:
        acc = accuracy_leg(tmp, args.accuracy_particles, args.accuracy_turns, args.accuracy_tiers.split(","),
                           chunk_turns=args.accuracy_chunk, grid=_grid_arg(args.accuracy_grid),
                           sort_every=args.accuracy_sort)
        print(json.dumps({"accuracy_long": acc}), flush=True)
        return
    from pyffa_amd import workloads as W
    M = W.bunch("stable", n=args.particles, seed=0)
    if args.only_surrogate:
        print(json.dumps({"surrogate": surrogate_leg(M, tmp, 1, args.surrogate_steps, args.train_iters,
                                                     args.surrogate_zt)}), flush=True)
        return
    gp = machines(2, tmp)
    n = M.shape[0]
    log(f"inputs ready: {n} particles")
    B32, rk32, el32, S32, k32 = run_tier("fp32", M, gp, args.warmup, args.steps)
    log(f"fp32 tier {el32:.3f} s, kernels {k32}")
    del B32, rk32
    ps = n * args.steps * PUSHES
    out = {
        "metric": "particle-steps/s w/ 2.5D SC; surrogate+FP32 vs FP64 map gather at fixed error",
        "value": ps / el32, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el32 / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 state / f32 field", "data": "synthetic",
        "config": {"workload": f"{n / 1e6:g}M protons (1e13 total), 12-sector scaling FFAG r0=8.5 m k=7.6 "
                               "(D=-1.4 vertically stable variant), order-2 Taylor maps, aperture r 8.0-9.3 m "
                               "|z|<5 cm, 2.5D SC (ns,nx,nz)=(512,256,256) every 20 steps, 2 RF gaps 100 kV, "
                               "Boris dt=0.05 ns; FP32-field tier; one step = 1000 pushes, 50 SC solves, "
                               "5 azimuth sorts, 1 stop check",
                   "particles": n, "sc_grid": list(GRID[:3]), "sc_step": SC_STEP, "pushes_per_step": PUSHES},
    }
    if not args.no_fp64:
        k = args.fp64_steps or args.steps
        ps64 = n * k * PUSHES
        B64, rk64, el64, S64, k64 = run_tier("fp64", M, gp, args.warmup, k)
        del B64, rk64
        log(f"fp64 path (input order) {el64:.3f} s, kernels {k64}")
        B64s, rk64s, el64s, S64s, k64s = run_tier("fp64", M, gp, args.warmup, k, sort_every=SORT_EVERY)
        del B64s, rk64s
        log(f"fp64 path (azimuth-ordered) {el64s:.3f} s, kernels {k64s}")
        B64f, rk64f, el64f, S64f, k64f = run_tier("fp64", M, gp, args.warmup, k, sort_every=SORT_EVERY,
                                                  fuse_fp64=True)
        del B64f, rk64f
        log(f"fp64 path (azimuth-ordered, steps fused) {el64f:.3f} s, kernels {k64f}")
        same = all(bool((S64s[key] == S64f[key]).all()) for key in S64s)
        log(f"fp64 fused steps bit-identical to one launch per step: {same}")
        out["fp64_value"] = ps64 / el64
        out["fp64_sorted_value"] = ps64 / el64s
        out["fp64_fused_sorted_value"] = ps64 / el64f
        out["speedup_fp32_vs_fp64_fused_sorted"] = (ps / el32) / (ps64 / el64f)
        out["fp64_fused_sorted_kernel_ms"] = k64f
        out["fp64_fused_bit_identical"] = same
        out["speedup_fp32_vs_fp64"] = (ps / el32) / (ps64 / el64)
        out["speedup_fp32_vs_fp64_sorted"] = (ps / el32) / (ps64 / el64s)
        out["fp64_kernel_ms"] = k64
        out["fp64_sorted_kernel_ms"] = k64s
        if k == args.steps:
            out["accuracy_fp32_vs_fp64"] = accuracy(S32, S64, (args.warmup + args.steps) * PUSHES)
        out["field_error_fp32_vs_fp64"] = field_error_leg(gp)
        log(f"field error fp32 vs fp64 {out['field_error_fp32_vs_fp64']}")
    # roofline of the dominant kernel of the headline tier: the fused 20-step
    # push, which also applies the preceding SC solve's gather and endpoint
    # roll.  Algorithmic HBM bytes per particle per launch: FP64 state read +
    # write (7 x 8 x 2), flag byte read + write, FP32 SC endpoints (the old
    # pre-step field read, the new one written: 2 x 3 x 4), the solve's
    # per-particle gather inputs (slice index 4, x = r - r0 8, pr0 8).  The
    # field-map and field-cube reads (cache-resident tables shared by many
    # particles) are not counted.
    bytes_per_particle = 7 * 8 * 2 + 2 + 2 * 3 * 4 + 4 + 8 + 8
    algo = n * bytes_per_particle
    ach = algo / (k32["push"] * 1e-3) / 1e9
    # the fused push keeps the state in registers for 20 steps, so it is
    # bound by vector-ALU issue rather than HBM: the measured VALU busy
    # fraction (rocprofv3 VALUBusy) is reported beside the HBM fraction
    out["roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": ach / HBM_PEAK_GBS, "traffic": measured_traffic(PUSH_KERNEL),
                       "algorithmic_bytes": algo, "kernel": "k_boris_fused<float> (20 steps + the solve's gather)",
                       "kernel_ms": k32["push"], "sc_solve_ms": k32["sc"], "valu": measured_valu(PUSH_KERNEL)}
    if not args.no_surrogate:
        out["surrogate"] = surrogate_leg(M, tmp, 1, args.surrogate_steps, args.train_iters, args.surrogate_zt)
    if not args.no_mapgen:
        out["map_generator"] = mapgen_leg(cpu=not args.no_cpu)
    if args.tune_turns > 0:
        out["tune_shift"] = tune_leg(tmp, args.tune_particles, args.tune_turns, args.tune_tracked,
                                     progress=args.tune_progress, grid=_grid_arg(args.tune_grid),
                                     sort_every=args.tune_sort)
    if args.cycle_steps > 0:
        out["spiral_cycle"] = cycle_leg(tmp, args.cycle_particles, args.cycle_steps)
    if args.accuracy_turns > 0:
        out["accuracy_long"] = accuracy_leg(tmp, args.accuracy_particles, args.accuracy_turns,
                                            args.accuracy_tiers.split(","), chunk_turns=args.accuracy_chunk,
                                            grid=_grid_arg(args.accuracy_grid), sort_every=args.accuracy_sort)
    if not args.no_cpu:
        cb = cpu_baseline(M, gp, min(args.cpu_particles, n), args.cpu_blocks, n)
        log(f"cpu baseline {cb}")
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
