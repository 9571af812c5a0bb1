"""Tracking loops of the hot path.

``FFAG_RungeKutta.rk4_solve_dt_bunch3_withSC_boris`` and
``rk4_solve_dt_bunch3_withSC`` keep the reference's signatures
(src/FFAG_track.py:297, :65) and control flow: per step, injection
(src/FFAG_Bunch.py:265-303), the 2.5D SC solve every ``SC_step`` steps with
the Pre/PrePre endpoint roll (src/FFAG_track.py:356-382), the linear SC
extrapolation (:449-478), the push, the detectors, a progress line and the
stop check every 1000 steps (:518-577).  The per-particle work runs in the
HIP kernels of libffa_hip.so; the host computes the per-step scalars
(f_RF, V0, extrapolation factor, injection range) and issues launches on the
torch stream, synchronising only for the progress line and the stop check.

``LocalParameters`` accepts, besides the reference keys, ``precision``
('fp64' default, 'fp32' for the FP32-field tier, which runs the steps
between two SC solves as one fused launch), ``verbose`` (progress lines),
``sort_every`` (steps between azimuth sorts of the particles: 200 for the
FP32 tier, 0 (never, the input order) for the FP64 path by default),
``start_step`` (continue a run: the step index the loop starts from, so the
SC schedule and endpoint history carry on; ``max_stepsN`` stays the absolute
end), ``max_block`` (cap on the steps issued per launch block) and
``fuse_fp64`` (FP64 Boris path: run each block of steps as one launch with
the state in registers, ``ffa_push_boris_f64_fused``; bit-identical to the
default one launch per step), ``defer_sc_gather`` (FP32 tier without the
surrogate, default True: each SC solve stops after the field and the push
launch that follows gathers the kicks, ``ffa_sc_solve_field`` +
``ffa_push_boris_fused_f32_sc``; bit-identical to the separate gather) and
``kernel_timer`` (an object whose ``begin(name)`` / ``end(name, launches)``
the loop calls around each push call, name ``"push"``, with the number of
kernel launches it issued (1 for a fused FP32 block, one per step on the FP64
path), and around each SC solve, ``"sc"``, on the host thread that issues
them: a caller can record stream events there to time the launches inside a
run).
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib
from .FFAG_Field import build_machine
from .FFAG_dump import Dumps


def FunctionForAccelerationBunch_dt(*args, **kwargs):
    """Signature placeholder for the reference's RK4 right-hand side
    (src/FFAG_track.py:1743-1801): the RK4 loop here evaluates it inside the
    ffa_push_rk4_f64 kernel, so the callable passed as ``func`` is ignored."""
    raise NotImplementedError("evaluated on the GPU by ffa_push_rk4_f64")


class FFAG_RungeKutta:
    def __init__(self):
        self.running = True
        self.ctx = None
        self.last_run = {}

    # --------------------------------------------------------------- set-up
    def _setup(self, BunchObj, GlobalParameters, precision):
        torch = BunchObj.torch
        if (precision == "fp32") != (BunchObj.precision == "fp32"):
            raise ValueError(f"bunch precision {BunchObj.precision!r} does not match LocalParameters "
                             f"precision {precision!r}")
        stream = torch.cuda.current_stream(BunchObj.device).cuda_stream
        if self.ctx is None:
            self.ctx = _lib.Context(BunchObj.device.index or 0, stream)
        else:
            self.ctx.check(self.ctx.lib.ffa_set_stream(self.ctx.h, _lib._vp(stream)), "ffa_set_stream")
        dt = torch.float32 if precision == "fp32" else torch.float64
        tay = GlobalParameters.Bmap.device_taylor(dt, BunchObj.device)
        self._taylor = tay
        m = build_machine(GlobalParameters, tay, precision == "fp32")
        self.ctx.check(self.ctx.lib.ffa_set_machine(self.ctx.h, m), "ffa_set_machine")
        self._phi_out = torch.zeros(2, dtype=torch.float64, device=BunchObj.device)

    @staticmethod
    def _rf(GlobalParameters, t):
        em = GlobalParameters.Emap
        if em is None:
            return 0.0, 0.0
        return em.Interpolation1D_freq_curve(t)

    def _stop(self, BunchObj, t, max_fi, max_time):
        lib = self.ctx.lib
        self.ctx.check(lib.ffa_reduce_phi(self.ctx.h, BunchObj.particles(), _lib._vp(self._phi_out.data_ptr())),
                       "ffa_reduce_phi")
        s, c = self._phi_out.cpu().tolist()
        if c > 0 and abs(s / c) > max_fi:
            return True
        return t > max_time

    def _progress(self, BunchObj, steps_i, t_post):
        fl = BunchObj.flags
        idx = (fl & _lib.F_ALIVE).nonzero()[:, 0]
        BunchObj.UpdateParticleNumGlobal()
        clock = time.strftime("%m-%d %H:%M:%S", time.localtime())
        if idx.numel() > 0:
            # the first alive particle in input order (the FP32 tier reorders)
            i = int(idx[int(BunchObj.pid[idx].argmin())])
            r = float(BunchObj.r[i])
            fi = float(BunchObj.fi[i])
            print(f"current step={steps_i:.0f}, r ~= {r:.2f} m, turn ~= {fi / np.pi / 2:.2f}, "
                  f"remain particles: {BunchObj.SurvivedNum}. tracking time: {t_post * 1e6:.2f} us,",
                  f"clock: {clock}", flush=True)
        else:
            print(f"current step={steps_i:.0f}, r ~= nan, turn ~= nan, remain particles: {BunchObj.SurvivedNum}. "
                  f"tracking time: {t_post * 1e6:.2f} us. ", f"clock: {clock}", flush=True)

    # ------------------------------------------------------------------ loops
    def rk4_solve_dt_bunch3_withSC_boris(self, func, t_start, BunchObj, GlobalParameters, LocalParameters):
        return self._run(t_start, BunchObj, GlobalParameters, LocalParameters, solver="boris")

    def rk4_solve_dt_bunch3_withSC(self, func, t_start, BunchObj, GlobalParameters, LocalParameters):
        return self._run(t_start, BunchObj, GlobalParameters, LocalParameters, solver="rk4")

    def _run(self, t_start, BunchObj, GlobalParameters, LocalParameters, solver):
        stop = LocalParameters["stop_condition"]
        step_t = float(LocalParameters["time_step"])
        enable_sc = bool(LocalParameters.get("enable_SC", False))
        sc_type = LocalParameters.get("SC_type", 1)
        sc_step = int(LocalParameters.get("SC_step", 1))
        dumps = LocalParameters.get("step_dumps")
        if dumps is None:
            dumps = Dumps()
        precision = LocalParameters.get("precision", BunchObj.precision)
        if enable_sc and sc_type != 1:
            raise TypeError("only SC_type=1 (2.5D slice-by-slice) is implemented on the GPU path")
        if solver == "rk4" and precision != "fp64":
            raise ValueError("the RK4 pusher runs in FP64 only")
        if sc_step < 1:
            raise ValueError("SC_step must be >= 1")
        max_steps = int(stop["max_stepsN"])
        max_fi = stop["max_turn"] * np.pi * 2
        max_time = stop["max_time"]
        surrogate = LocalParameters.get("surrogate")
        if surrogate is not None and not (precision == "fp32" and solver == "boris"):
            raise ValueError("the field surrogate drives the FP32 Boris tier only")
        self._setup(BunchObj, GlobalParameters, precision)
        ctx = self.ctx
        lib = ctx.lib
        if surrogate is not None:
            surrogate.enable(ctx)
        try:
            return self._loop(t_start, BunchObj, GlobalParameters, LocalParameters, solver, precision, dumps,
                              enable_sc, sc_step, step_t, max_steps, max_fi, max_time)
        finally:
            if surrogate is not None:
                surrogate.disable(ctx)

    def _loop(self, t_start, BunchObj, GlobalParameters, LocalParameters, solver, precision, dumps, enable_sc,
              sc_step, step_t, max_steps, max_fi, max_time):
        ctx = self.ctx
        lib = ctx.lib
        verbose = LocalParameters.get("verbose", True)
        has_dumps = len(getattr(dumps, "dumps", [1])) > 0
        needs_step = getattr(dumps, "needs_step", None)
        # FP32 tier: reorder particles by azimuth every sort_every steps (map
        # and SC-grid locality; at 8M the 20-step push goes from 1.27 ms
        # after a sort to 1.58 ms 1000 steps later, and a sort costs about
        # 1 ms, so every 200 steps is the measured optimum); the FP64 path
        # keeps the input order
        sort_every = int(LocalParameters.get("sort_every", 200 if precision == "fp32" else 0))
        t_pre = t_start
        steps_i = int(LocalParameters.get("start_step", 0))
        max_block = int(LocalParameters.get("max_block", _MAX_BLOCK))
        fused = precision == "fp32" and solver == "boris"
        fuse64 = solver == "boris" and not fused and bool(LocalParameters.get("fuse_fp64", False))
        # FP32 tier without the surrogate: each solve stops after the field
        # and its gather runs inside the push launch that follows
        # (bit-identical; one pass over the particles fewer per solve)
        defer_gather = (fused and LocalParameters.get("surrogate") is None
                        and bool(LocalParameters.get("defer_sc_gather", True)))
        timer = LocalParameters.get("kernel_timer")
        n_solves = 0
        while steps_i < max_steps:
            # one block: the steps up to the next SC solve / 1000-step check /
            # end; a block is one fused launch (FP32 tier) or one C-side
            # sequence of per-step launches (FP64 path).  A step that a
            # detector examines runs as a block of its own, with the pre-step
            # state recorded; the steps before it form the previous block.
            nb = 1000 - steps_i % 1000
            if enable_sc:
                nb = min(nb, sc_step - steps_i % sc_step)
            if sort_every > 0:
                nb = min(nb, sort_every - steps_i % sort_every)
            nb = min(nb, max_steps - steps_i, max(1, min(max_block, _MAX_BLOCK)))
            examine = False
            if has_dumps:
                j = 0
                if needs_step is not None:
                    tt = t_pre
                    while j < nb and not needs_step(tt, tt + step_t):
                        tt = tt + step_t
                        j += 1
                if j == 0:
                    nb, examine = 1, True
                else:
                    nb = j
            b0, e0 = BunchObj.injection_range(t_pre, step_t)
            if e0 > b0:
                ctx.check(lib.ffa_inject(ctx.h, BunchObj.particles(), b0, e0), "ffa_inject")
                BunchObj._account_injection(b0, e0)
            if sort_every > 0 and steps_i % sort_every == 0:
                BunchObj.sort_by_azimuth(ctx)
            if enable_sc and steps_i % sc_step == 0:
                s_last = ((steps_i - 1) % sc_step) / sc_step if steps_i - 1 >= sc_step else 0.0
                if timer is not None:
                    timer.begin("sc")
                BunchObj.sc_solve_device(ctx, first=(steps_i == 0), s_last=s_last, defer_gather=defer_gather)
                if timer is not None:
                    timer.end("sc", 1)
                n_solves += 1
            steps = (_lib.FfaStep * nb)()
            inj = (_lib._i64 * (2 * nb))()
            tt = t_pre
            for j in range(nb):
                si = steps_i + j
                if j > 0:
                    b, e = BunchObj.injection_range(tt, step_t)
                    if e > b:
                        inj[2 * j], inj[2 * j + 1] = b, e
                        BunchObj._account_injection(b, e)
                st = steps[j]
                st.t, st.dt = tt, step_t
                st.sc_s = (si % sc_step) / sc_step if si >= sc_step else 0.0
                if solver == "boris":
                    f, v0 = self._rf(GlobalParameters, tt)
                    st.dRF = f * np.pi * 2 * step_t
                    st.v0 = v0
                else:
                    for q, tq in enumerate((tt, tt + 0.5 * step_t, tt + step_t)):
                        fq, vq = self._rf(GlobalParameters, tq)
                        st.rk_rate[q] = fq * np.pi * 2
                        st.rk_v0[q] = vq
                tt = tt + step_t
            if examine:
                BunchObj.UpdatePreSteps()
            if timer is not None:
                timer.begin("push")
            if fused:
                BunchObj.push_fused_f32(ctx, steps, inj, nb)
            elif fuse64:
                ctx.check(lib.ffa_push_boris_f64_fused(ctx.h, BunchObj.particles(), steps, inj, nb),
                          "ffa_push_boris_f64_fused")
            else:
                ctx.check(lib.ffa_push_f64_multi(ctx.h, BunchObj.particles(), steps, inj, nb,
                                                 1 if solver == "rk4" else 0), "ffa_push_f64_multi")
            if timer is not None:
                timer.end("push", 1 if fused or fuse64 else nb)
            BunchObj.last_sc_s = steps[nb - 1].sc_s
            BunchObj.UpdatePostSteps()
            if examine:
                dumps.check_and_dump(t_pre, tt, BunchObj, data_type="Boris" if solver == "boris" else "RK4")
            t_pre = tt
            first_step = steps_i
            steps_i += nb
            # blocks never cross a multiple of 1000, so the reference's
            # "steps_i % 1000 == 0" progress line falls on a block's first step
            if verbose and first_step % 1000 == 0:
                self._progress(BunchObj, first_step, t_pre)
            if steps_i % 1000 == 0:
                if self._stop(BunchObj, t_pre, max_fi, max_time):
                    break
        self.last_run = dict(steps=steps_i, t=t_pre, sc_solves=n_solves)
        BunchObj.invalidate_host()
        return 0


_MAX_BLOCK = 1000
