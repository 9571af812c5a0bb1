"""Throughput benchmark of the FFAG tracking hot path on one MI355X.

Workload: 12-sector scaling FFAG (r0 = 8.5 m, k = 7.6, 300 MeV protons,
order-2 Taylor maps on an r 8.0-9.3 m x one-sector grid), 1M macro-particles
(seeded Gaussian bunch, uniform azimuth), 2.5D space charge on a
(ns, nx, nz) = (256, 128, 128) grid at +-4 sigma every 20 push steps, two RF
gaps 180 degrees apart at 100 kV with a linear frequency ramp, Boris pusher,
dt = 0.05 ns.  One bench step = 20 push steps + one SC solve.

The headline value is the FP32-field tier (FP64 state, fused steps); the FP64
map-gather path is timed on the same workload for the speed-up, and the
accuracy of the FP32 tier is reported against it (centroid, rms emittance).
The CPU baseline is the OpenMP FP64 oracle on the box's host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_PART = 1_000_000
GRID = (256, 128, 128, 4.0)      # (ns, nx, nz, sigma_ratio)
SC_STEP = 20
DT = 0.05e-9
HBM_PEAK_GBS = 8000.0


def build_inputs(n, tmpdir):
    import numpy as np
    from tests import cases
    mc, gp = cases.machines("stable", max_order=2, rf=dict(v0=100e3), aperture=cases.APERTURE, tmpdir=tmpdir)
    M = cases.bunch("stable", n=n, seed=0)
    return mc, gp, M


def snapshot(B):
    return {k: getattr(B, k).clone() for k in ("r", "vr", "z", "vz", "fidot", "flags")}


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


def run_tier(precision, M, gp, warmup, steps):
    import torch
    from pyffa_amd.FFAG_Bunch import FFAG_Bunch
    from pyffa_amd.FFAG_track import FFAG_RungeKutta
    from tests import cases
    B = FFAG_Bunch(M, marcosize=1e13 / M.shape[0], sc_grid_size=GRID, BunchType="Boris", precision=precision)
    cases.set_bunch_seo(B)
    rk = FFAG_RungeKutta()

    def lp(first, last):
        return dict(stop_condition=dict(max_stepsN=last * SC_STEP, max_turn=1e9, max_time=1.0),
                    time_step=DT, step_dumps=None, enable_SC=True, SC_type=1, SC_step=SC_STEP,
                    precision=precision, verbose=False, start_step=first * SC_STEP)
    t = 0.0
    if warmup:
        rk.rk4_solve_dt_bunch3_withSC_boris(None, t, B, gp, lp(0, warmup))
        t = rk.last_run["t"]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rk.rk4_solve_dt_bunch3_withSC_boris(None, t, B, gp, lp(warmup, warmup + steps))
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return B, rk, el, snapshot(B)


def kernel_time_ms(rk, B, gp, precision, reps=10):
    """Average duration of the dominant kernel launch (fused 20-step push for
    the FP32 tier) and of one SC solve, with HIP events on the stream the
    kernels run on."""
    import torch
    from pyffa_amd import _lib
    import numpy as np
    ctx = rk.ctx
    stream = torch.cuda.current_stream()
    steps = (_lib.FfaStep * SC_STEP)()
    inj = (_lib._i64 * (2 * SC_STEP))()
    t = rk.last_run["t"]
    for j in range(SC_STEP):
        f, v0 = gp.Emap.Interpolation1D_freq_curve(t + j * DT)
        steps[j].t, steps[j].dt, steps[j].dRF, steps[j].v0 = t + j * DT, DT, f * np.pi * 2 * DT, v0
        steps[j].sc_s = j / SC_STEP
    res = {}
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ts = []
    for _ in range(reps):
        ev[0].record(stream)
        if precision == "fp32":
            ctx.check(ctx.lib.ffa_push_boris_fused_f32(ctx.h, B.particles(), steps, inj, SC_STEP), "fused")
        else:
            ctx.check(ctx.lib.ffa_push_boris_f64(ctx.h, B.particles(), steps[0]), "f64")
        ev[1].record(stream)
        ev[1].synchronize()
        ts.append(ev[0].elapsed_time(ev[1]))
    res["push"] = float(np.mean(ts[1:]))
    ts = []
    for _ in range(reps):
        ev[2].record(stream)
        B.sc_solve_device(ctx, first=False, s_last=0.95)
        ev[3].record(stream)
        ev[3].synchronize()
        ts.append(ev[2].elapsed_time(ev[3]))
    res["sc"] = float(np.mean(ts[1:]))
    return res


def cpu_baseline(M, mc, n_sample, macro_steps):
    import numpy as np
    from oracle import oracle as O
    from tests import cases
    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    O.lib().ora_set_threads(threads)
    S = M[:n_sample].copy()
    tr = O.Tracker(mc, S, enable_sc=True, sc_step=SC_STEP, grid=GRID, marcosize=1e13 / n_sample,
                   seo=cases.oracle_seo())
    t0 = time.perf_counter()
    tr.run(0.0, DT, macro_steps * SC_STEP)
    el = time.perf_counter() - t0
    return dict(value=n_sample * macro_steps * SC_STEP / el, unit="particle-steps/s", cores=threads, kind="port",
                sample=f"{n_sample} particles x {macro_steps * SC_STEP} Boris steps with "
                       f"{macro_steps} SC solves (FP64 OpenMP oracle), {el:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300, help="timed bench steps (20 pushes + 1 SC solve each)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--particles", type=int, default=N_PART)
    ap.add_argument("--cpu-particles", type=int, default=1_000_000)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import tempfile
    import torch
    tmp = tempfile.mkdtemp()
    mc, gp, M = build_inputs(args.particles, tmp)
    print(f"[bench] inputs ready: {M.shape[0]} particles", file=sys.stderr, flush=True)
    B32, rk32, el32, S32 = run_tier("fp32", M, gp, args.warmup, args.steps)
    k32 = kernel_time_ms(rk32, B32, gp, "fp32")
    print(f"[bench] fp32 tier {el32:.3f} s, kernels {k32}", file=sys.stderr, flush=True)
    n_alive = M.shape[0]
    ps = n_alive * args.steps * SC_STEP
    out = {
        "metric": "particle-steps/s w/ 2.5D SC; surrogate+FP32 vs FP64 map gather at fixed error",
        "value": ps / el32, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el32 / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 state / f32 field", "data": "synthetic",
        "config": {"workload": "cfg1: 1M protons (1e13 total), 12-sector scaling FFAG r0=8.5 m k=7.6 "
                               "(D=-1.4 vertically stable variant), order-2 Taylor maps, aperture r 8.0-9.3 m |z|<5 cm, 2.5D SC (ns,nx,nz)=(256,128,128) every 20 steps, 2 RF gaps "
                               "100 kV, Boris dt=0.05 ns; FP32-field tier",
                   "particles": n_alive, "sc_grid": list(GRID[:3]), "sc_step": SC_STEP},
    }
    if not args.no_fp64:
        B64, rk64, el64, S64 = run_tier("fp64", M, gp, args.warmup, args.steps)
        both = ((S32["flags"] & 2) != 0) & ((S64["flags"] & 2) != 0)
        m32, m64 = moments(S32, both), moments(S64, both)
        k64 = kernel_time_ms(rk64, B64, gp, "fp64")
        print(f"[bench] fp64 path {el64:.3f} s, kernels {k64}", file=sys.stderr, flush=True)
        out["fp64_value"] = ps / el64
        out["speedup_fp32_vs_fp64"] = el64 / el32
        out["accuracy_fp32_vs_fp64"] = {
            "centroid_r_abs_m": abs(m32["r_mean"] - m64["r_mean"]),
            "centroid_z_abs_m": abs(m32["z_mean"] - m64["z_mean"]),
            "emit_r_rel": abs(m32["emit_r"] / m64["emit_r"] - 1.0),
            "emit_z_rel": abs(m32["emit_z"] / m64["emit_z"] - 1.0),
            "alive_fp32": int(((S32["flags"] & 2) != 0).sum()), "alive_fp64": int(((S64["flags"] & 2) != 0).sum()),
            "compared": m32["alive"], "after_steps": (args.warmup + args.steps) * SC_STEP}
        out["fp64_kernel_ms"] = k64
    # roofline of the dominant kernel of the headline tier: the fused 20-step push.
    # Algorithmic HBM bytes per particle per launch: FP64 state read + write
    # (7 x 8 x 2), flag byte read + write, FP32 SC endpoints Pre/PrePre (6 x 4).
    bytes_per_particle = 7 * 8 * 2 + 2 + 6 * 4
    traffic = n_alive * bytes_per_particle
    ach = traffic / (k32["push"] * 1e-3) / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": ach / HBM_PEAK_GBS, "traffic": None, "kernel": "k_boris_fused<float> (20 steps)",
                       "kernel_ms": k32["push"], "sc_solve_ms": k32["sc"]}
    if not args.no_cpu:
        cb = cpu_baseline(M, mc, min(args.cpu_particles, M.shape[0]), args.cpu_steps)
        print(f"[bench] cpu baseline {cb}", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
