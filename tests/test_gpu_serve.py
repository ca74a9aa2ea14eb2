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
