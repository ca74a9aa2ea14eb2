"""Continuous-batching serving loop (mojoxm_amd.serve): every request's
result is the same, bit for bit, served in the mix as solved alone, and
the loop respects its HBM byte budget."""
import numpy as np
import pytest

import mojoxm_amd as M
from mojoxm_amd import problems, serve

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_FP32_MFMA])
def test_serving_results_equal_alone(precision):
    reqs = problems.serving_stream(count=8, rate=200.0, seed=4, n_min=8, n_max=20, t_range=(0.004, 0.02))
    states = {r.index: serve.initial_state(r) for r in reqs}
    if precision == M.PREC_FP32_MFMA:
        states = {k: v.astype(np.float32).astype(np.float64) for k, v in states.items()}
    rep = serve.serve(reqs, precision, chunk=5, keep_results=True, states=states)
    secs, alone = serve.one_at_a_time(reqs, precision, states=states, keep_results=True)
    assert rep.solves == len(reqs)
    for r in reqs:
        assert np.isfinite(rep.results[r.index]).all()
        assert np.array_equal(rep.results[r.index], alone[r.index]), (r.index, r.kind)
    assert len(serve.fifo_latencies(reqs, secs)) == len(reqs)


def test_serving_budget_limits_memory():
    reqs = problems.serving_stream(count=6, rate=1000.0, seed=7, n_min=10, n_max=16, t_range=(0.002, 0.005))
    fps = [M.batch.footprint(r.n, r.L, r.physics, r.P, M.PREC_FP64, r.limiter, r.bcs) for r in reqs]
    budget = 2 * max(fps)
    rep = serve.serve(reqs, M.PREC_FP64, budget=budget, chunk=2)
    assert rep.solves == len(reqs)
    assert rep.peak_bytes <= budget


def test_serving_evicts_a_failed_request():
    """A request whose initial state holds a NaN fails alone: it is switched
    off on the device, reported in ServeReport.failed and evicted (its
    memory returns to the budget, which holds only a few requests at a
    time), and every other request's result equals its solve alone."""
    reqs = problems.serving_stream(count=6, rate=1000.0, seed=11, n_min=8, n_max=14, t_range=(0.004, 0.01))
    states = {r.index: serve.initial_state(r) for r in reqs}
    poison = next(r for r in reqs if r.kind != "sod")  # the Euler rescue would floor a NaN density
    states[poison.index] = states[poison.index].copy()
    states[poison.index][3, 2, 0] = np.nan
    fps = [M.batch.footprint(r.n, r.L, r.physics, r.P, M.PREC_FP64, r.limiter, r.bcs) for r in reqs]
    budget = 6 * max(fps)
    rep = serve.serve(reqs, M.PREC_FP64, budget=budget, chunk=3, keep_results=True, states=states)
    assert rep.failed == [poison.index]
    assert rep.solves == len(reqs) - 1 and rep.peak_bytes <= budget
    good = [r for r in reqs if r is not poison]
    _, alone = serve.one_at_a_time(good, M.PREC_FP64, states=states, keep_results=True)
    for r in good:
        assert np.array_equal(rep.results[r.index], alone[r.index]), (r.index, r.kind)
