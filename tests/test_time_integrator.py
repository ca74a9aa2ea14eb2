"""Host logic of the time loops (src/time_integrator.mojo:61-179): step
clamping to frame boundaries and to t_final, the frame hooks, and the
dt rules.  A stub solver records what the loop asks of it; no device."""
import pytest

from mojoxm_amd.time_integrator import (apriori_dt, fixed_steps, run_ssprk3_loop,
                                        run_ssprk3_loop_with_diagnostics)


class _Solver:
    P = 2

    def __init__(self):
        self.dts, self.events = [], []

    def step_ssprk3(self, dt):
        self.dts.append(dt)
        self.events.append(("step", dt))

    def synchronize(self):
        self.events.append(("sync",))


class _Hook:
    def __init__(self, solver, tag):
        self.solver, self.tag, self.times = solver, tag, []

    def write(self, t):
        self.times.append(t)
        self.solver.events.append((self.tag, t))

    record = write


def test_steps_clamp_to_frames_and_end():
    s = _Solver()
    seen = []
    r = run_ssprk3_loop(s, 0.3, 1.0, num_frames=2, on_frame=lambda so, t: seen.append(t))
    assert s.dts == pytest.approx([0.3, 0.2, 0.3, 0.2])
    assert seen == pytest.approx([0.5, 1.0])
    assert r.total_steps == 4 and r.t_final == pytest.approx(1.0)
    assert r.frame_write_sec >= 0.0 and r.wall_sec >= r.step_loop_sec


def test_clock_round_off_step():
    """The loop runs while t < t_final, as the reference's does: when the
    summed dt falls short of t_final by round-off, one more step of that
    round-off's size lands on t_final exactly."""
    s = _Solver()
    n, dt = fixed_steps(1.0, 0.07)
    r = run_ssprk3_loop(s, dt, 1.0)
    assert n == 15 and r.total_steps in (15, 16)
    assert all(x == dt for x in s.dts[:15]) and all(x < 1e-12 for x in s.dts[15:])
    assert r.t_final == 1.0


def test_with_diagnostics_writes_t0_and_every_frame():
    s = _Solver()
    w, d = _Hook(s, "frame"), _Hook(s, "diag")
    r = run_ssprk3_loop_with_diagnostics(s, w, d, 0.1, 0.6, num_frames=3)
    assert w.times == pytest.approx([0.0, 0.2, 0.4, 0.6])
    assert d.times == w.times
    assert r.total_steps == 6
    # t = 0 before any step; each later frame after a synchronize and
    # before the next step
    assert s.events[:2] == [("frame", 0.0), ("diag", 0.0)]
    i = s.events.index(("frame", w.times[1]))
    assert s.events[i - 1] == ("sync",) and s.events[i + 1][0] == "diag" and s.events[i + 2][0] == "step"
    assert s.events[-1] == ("sync",)


def test_with_diagnostics_accepts_missing_hooks():
    s = _Solver()
    d = _Hook(s, "diag")
    r = run_ssprk3_loop_with_diagnostics(s, None, d, 0.25, 1.0, num_frames=1)
    assert d.times == pytest.approx([0.0, 1.0]) and r.total_steps == 4


def test_apriori_dt():
    # dt = CFL h / (s (2P+1))
    assert apriori_dt(0.5, 0.1, 2.0, 3) == pytest.approx(0.5 * 0.1 / (2.0 * 7))
