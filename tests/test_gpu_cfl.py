"""The device CFL reduction (mxm_dg_max_wave_speed, mxm_batch_max_wave_speed)
against the oracle's largest-signal-speed rule (oracle/cfl.c) on seeded
states of every physics, in both arithmetic forms (the speed is evaluated in
FP64 from the state, so the FP32 form, holding an FP32-representable state,
gives the same value up to the FP64 square root's last bits); the batch
against the single solver; NaN propagation; the cfl_dt helper."""
import math

import numpy as np
import pytest

import mojoxm_amd as M
from mojoxm_amd import physics as ph
from mojoxm_amd.batch import Batch
from mojoxm_amd.time_integrator import apriori_dt, cfl_dt

from test_gpu_parity import random_state

pytestmark = pytest.mark.gpu

CASES = [
    ((3, 2, 2), ph.Advection(1.0, -0.5, 0.3), 2),
    ((3, 2, 2), ph.Euler(flux_type=ph.FLUX_HLLEC), 3),
    ((6, 5), ph.Euler(flux_type=ph.FLUX_ROE), 2),
    ((6, 5), ph.Euler2D(hllc=True), 3),
    ((6, 5), ph.ShallowWater2D(g=9.81), 2),
    ((6, 5), ph.MHDGLM2D(c_h=0.5, alpha_d=2.0), 3),
    ((6, 5), ph.Maxwell2D(c=1.7), 2),
    ((6, 5), ph.Advection2D(0.6, -0.8), 1),
    ((3, 2, 2), ph.ShallowWater(g=9.81), 2),
    ((3, 2, 2), ph.Maxwell(c=1.3), 2),
    ((3, 2, 2), ph.IdealMHD(c_h=0.5, alpha_d=1.0), 3),
    ((3, 2, 2), ph.FiveMomentTwoFluid(c_light=0.5, c_h=0.4), 2),
]


def box(n):
    return (1.0, 0.9, 1.1)[: len(n)]


def nodes(ndim, P):
    return (P + 1) * (P + 2) * (P + 3) // 6 if ndim == 3 else (P + 1) * (P + 2) // 2


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_FP32_MFMA])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_max_wave_speed_matches_oracle(oracle, case, precision):
    n, phys, P = CASES[case]
    K = int(np.prod(n)) * (6 if len(n) == 3 else 2)
    q = random_state(phys, K, nodes(len(n), P), np.random.default_rng(40 + case), len(n))
    q = q.astype(np.float32).astype(np.float64)
    want = oracle.max_wave_speed(phys.phys_id, phys.params(), q)
    with M.Solver(n, box(n), phys, P=P, precision=precision) as s:
        s.set_state(q)
        got = s.max_wave_speed()
        assert cfl_dt(s, 0.2, 0.1) == pytest.approx(apriori_dt(0.2, 0.1, got, P), rel=1e-15)
    assert got == pytest.approx(want, rel=1e-13)


def test_batch_equals_solver_and_nan():
    rng = np.random.default_rng(5)
    phys = ph.Euler(flux_type=ph.FLUX_HLLEC)
    qs = [random_state(phys, 2 * a * b, 10, rng, 2) for a, b in ((6, 5), (4, 7))]
    with Batch(M.PREC_FP64) as b:
        ids = [b.add(nn, (1.0, 1.0), phys, 3, q, 1e-3, 2) for nn, q in zip(((6, 5), (4, 7)), qs)]
        got = [b.max_wave_speed(i) for i in ids]
    for nn, q, g in zip(((6, 5), (4, 7)), qs, got):
        with M.Solver(nn, (1.0, 1.0), phys, P=3) as s:
            s.set_state(q)
            assert s.max_wave_speed() == g
    bad = qs[0].copy()
    bad[7, 3, 2] = np.nan
    with M.Solver((6, 5), (1.0, 1.0), phys, P=3) as s:
        s.set_state(bad)
        assert math.isnan(s.max_wave_speed())
