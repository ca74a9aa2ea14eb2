"""Launch configurations of the stage kernel (mxm_dg_autotune,
mxm_dg_launch_config, mxm_batch_autotune).

Every configuration is the same kernel source compiled under a different
occupancy target, so every configuration must produce the same bits; the
autotuner may then pick any of them without changing a result.  This is
checked for both arithmetic forms on each physics family, and the
autotuner is checked not to move the state it times on."""
import numpy as np
import pytest

import mojoxm_amd as M
from mojoxm_amd import physics as ph
from mojoxm_amd.batch import Batch
from mojoxm_amd.solver import launch_config_count

from test_gpu_parity import random_state

pytestmark = pytest.mark.gpu

CASES = [
    ((6, 5), (1.0, 1.0), ph.Euler(flux_type=ph.FLUX_HLLEC), 3),
    ((6, 5), (1.0, 1.0), ph.Euler(flux_type=ph.FLUX_ROE), 1),
    ((5, 4), (1.0, 1.0), ph.MHDGLM2D(c_h=1.2, alpha_d=1.2 / 0.18), 3),
    ((5, 4), (1.0, 1.0), ph.ShallowWater2D(), 4),
    ((5, 4), (1.0, 1.0), ph.Maxwell2D(), 2),
    ((2, 3, 2), (1.0, 1.0, 1.0), ph.Euler(flux_type=ph.FLUX_RUSANOV), 3),
    ((2, 2, 2), (1.0, 1.0, 1.0), ph.Advection(1.0, -0.5, 0.3), 2),
    ((2, 2, 2), (1.0, 1.0, 1.0), ph.Maxwell(c=1.2, J=(0.1, 0.0, -0.1)), 3),
    ((2, 2, 2), (1.0, 1.0, 1.0), ph.IdealMHD(c_h=1.2, alpha_d=2.0), 2),
    ((2, 2, 2), (1.0, 1.0, 1.0), ph.FiveMomentTwoFluid(c_light=2.0, c_h=1.5, alpha_d=0.5), 2),
]


def advance(n, L, phys, P, precision, q0, cfg, nsteps=3):
    with M.Solver(n, L, phys, P=P, precision=precision) as s:
        s.launch_config = cfg
        assert s.launch_config == cfg
        s.set_state(q0)
        s.step_ssprk3(0.2 / max(n) / (2 * P + 1) / 2.0, nsteps)
        return s.get_state()


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_FP32_MFMA])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_configs_bit_identical(case, precision):
    n, L, phys, P = CASES[case]
    ndim = len(n)
    K = int(np.prod(n)) * (6 if ndim == 3 else 2)
    NP = (P + 1) * (P + 2) * (P + 3) // 6 if ndim == 3 else (P + 1) * (P + 2) // 2
    q0 = random_state(phys, K, NP, np.random.default_rng(11 + case), ndim)
    if precision == M.PREC_FP32_MFMA:
        q0 = q0.astype(np.float32).astype(np.float64)
    ncfg = launch_config_count(precision)
    assert ncfg >= 2
    ref = advance(n, L, phys, P, precision, q0, 0)
    assert np.isfinite(ref).all()
    for cfg in range(1, ncfg):
        assert np.array_equal(advance(n, L, phys, P, precision, q0, cfg), ref), cfg


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_FP32_MFMA])
def test_autotune_keeps_state(precision):
    phys = ph.Euler(flux_type=ph.FLUX_HLLEC)
    n, L, P = (12, 10), (1.0, 1.0), 3
    q0 = random_state(phys, 240, 10, np.random.default_rng(2), 2)
    with M.Solver(n, L, phys, P=P, precision=precision) as s:
        s.set_state(q0)
        before = s.get_state()
        cfg, us = s.autotune(1e-3, reps=5)
        assert 0 <= cfg < launch_config_count(precision) and s.launch_config == cfg
        assert len(us) == launch_config_count(precision) and min(us) > 0.0 and us[cfg] == min(us)
        assert np.array_equal(s.get_state(), before)
    with pytest.raises(M.MxmError):
        with M.Solver(n, L, phys, P=P, precision=precision) as s:
            s.launch_config = launch_config_count(precision)


def test_batch_autotune_on_off_same_bits():
    phys = ph.Euler(flux_type=ph.FLUX_HLLEC)
    rng = np.random.default_rng(4)
    q = [random_state(phys, 2 * nn * nn, 10, rng, 2) for nn in (6, 9)]
    out = []
    for on in (True, False):
        with Batch(M.PREC_FP32_MFMA) as b:
            b.autotune(on)
            ids = [b.add((nn, nn), (1.0, 1.0), phys, 3, q[k].astype(np.float32).astype(np.float64), 1e-3, 4)
                   for k, nn in enumerate((6, 9))]
            b.step(4)
            out.append([b.get_state(i) for i in ids])
    for a, c in zip(*out):
        assert np.array_equal(a, c)
