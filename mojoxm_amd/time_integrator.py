"""Fixed-dt SSPRK3 loop with frame clamping (run_ssprk3_loop,
src/time_integrator.mojo:61-116), its variant that writes frames and
diagnostics rows (src/time_integrator.mojo:118-179) and the drivers' a-priori dt rule
(examples/advection_gaussian.mojo:126-131)."""
from __future__ import annotations

import math
import time
from dataclasses import dataclass


@dataclass
class TimeLoopResult:
    total_steps: int
    wall_sec: float
    step_loop_sec: float
    final_sync_sec: float
    t_final: float
    frame_write_sec: float = 0.0  # frame-boundary sync + on_frame


def apriori_dt(cfl: float, h: float, wave: float, P: int) -> float:
    """dt = CFL h / (wave (2P+1))."""
    return cfl * h / (wave * (2 * P + 1))


def cfl_dt(solver, cfl: float, h: float) -> float:
    """dt = CFL h / (s (2P+1)) with s the largest signal speed of the
    solver's current state, reduced on the device (Solver.max_wave_speed):
    the measured counterpart of apriori_dt."""
    return apriori_dt(cfl, h, solver.max_wave_speed(), solver.P)


def fixed_steps(t_final: float, dt_est: float):
    """steps = ceil(T/dt), dt = T/steps (the config convention)."""
    n = max(1, int(math.ceil(t_final / dt_est - 1e-12)))
    return n, t_final / n


def run_ssprk3_loop_with_diagnostics(solver, writer, diag, dt: float, t_final: float,
                                     num_frames: int = 1) -> TimeLoopResult:
    """run_ssprk3_loop that also writes a frame (writer.write(t), e.g.
    mojoxm_amd.vtu.FrameWriter) and a diagnostics row (diag.record(t), a
    DiagnosticsWriter) at t = 0 and at every frame boundary
    (src/time_integrator.mojo:118-179); either may be None."""
    def frame(_, t):
        if writer is not None:
            writer.write(t)
        if diag is not None:
            diag.record(t)
    frame(solver, 0.0)
    return run_ssprk3_loop(solver, dt, t_final, num_frames, on_frame=frame)


def run_ssprk3_loop(solver, dt: float, t_final: float, num_frames: int = 1, on_frame=None) -> TimeLoopResult:
    """Step to t_final, clamping each step to land on frame boundaries and on
    t_final; on_frame(solver, t) is called at each frame boundary, after a
    synchronize, and the time both take is frame_write_sec."""
    t = 0.0
    frame_dt = t_final / num_frames
    next_frame = frame_dt
    frame_id = 1
    steps = 0
    tloop = 0.0
    twrite = 0.0
    w0 = time.perf_counter()
    while t < t_final:
        step_dt = dt
        if t + step_dt > next_frame:
            step_dt = next_frame - t
        if step_dt <= 0.0:
            step_dt = dt
        if t + step_dt > t_final:
            step_dt = t_final - t
        s0 = time.perf_counter()
        solver.step_ssprk3(step_dt)
        tloop += time.perf_counter() - s0
        t += step_dt
        steps += 1
        if t >= next_frame - 1e-12 and frame_id < num_frames + 1:
            if on_frame is not None:
                f0 = time.perf_counter()
                solver.synchronize()
                on_frame(solver, t)
                twrite += time.perf_counter() - f0
            frame_id += 1
            next_frame += frame_dt
    s0 = time.perf_counter()
    solver.synchronize()
    s1 = time.perf_counter()
    return TimeLoopResult(steps, s1 - w0, tloop, s1 - s0, t, twrite)
