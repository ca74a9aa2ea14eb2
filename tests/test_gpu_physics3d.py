"""Parity of the 3D system stage kernels (shallow water, Maxwell, ideal MHD
with GLM cleaning, five-moment two-fluid; Rusanov fluxes) with the CPU FP64
oracle (oracle/physics3d.c) through the C ABI, on seeded states, every
boundary kind and P = 1..3 (P = 4, 5 on a smaller mesh).

Bars: plain FP64 relative L2 <= 1e-12 after two SSPRK3 steps; FP32-MFMA
relative L2 <= 1e-5 after one step from an FP32-representable state.  The
batched engine must give the single solver's bits for the same instance,
and the diagnostics must match the oracle's."""
import numpy as np
import pytest

import mojoxm_amd as M
from mojoxm_amd import physics as ph
from mojoxm_amd.batch import Batch

from test_gpu_parity import random_state, rel_l2

pytestmark = pytest.mark.gpu

PHYS = {
    "sw": ph.ShallowWater(g=9.81, h_min=1e-6, inflow=(1.1, 0.2, -0.1)),
    "maxwell": ph.Maxwell(c=1.3, J=(0.1, -0.2, 0.3), M=(0.05, 0.0, -0.1), inflow=(0.1, 0.0, -0.2, 0.0, 0.3, 0.1)),
    "mhd": ph.IdealMHD(gamma=5.0 / 3.0, c_h=1.7, alpha_d=2.0, inflow=(1.0, 0.1, 0.0, 0.0, 2.0, 0.3, 0.1, 0.0, 0.0)),
    "twofluid": ph.FiveMomentTwoFluid(c_light=2.0, c_h=1.5, alpha_d=0.5,
                                      inflow=(1.0, 0.0, 0.0, 0.0, 1.5, 1.0, 0.1, 0.0, 0.0, 1.5)
                                      + (0.0,) * 7),
}


def state(name, K, NP, rng):
    return random_state(PHYS[name], K, NP, rng, 3)


BCS = [None, [1, 1, 1, 1, 1, 1], [2, 3, 1, 2, 3, 0]]


def run_pair(oracle, name, P, n, bc, nsteps, precision, seed):
    phys = PHYS[name]
    L = (1.0, 0.8, 1.2)
    bcs = M.BoundaryConditions(*bc) if bc else None
    with M.Solver(n, L, phys, P=P, bcs=bcs, precision=precision) as s:
        q = state(name, s.num_elements, s.NP, np.random.default_rng(seed))
        if precision == M.PREC_FP32_MFMA:
            q = q.astype(np.float32).astype(np.float64)
        dt = 0.1 * min(L) / max(n) / (2 * P + 1) / 3.0
        s.set_state(q)
        s.step_ssprk3(dt, nsteps)
        got = s.get_state()
    om = oracle.Mesh(3, P, n, L, bc)
    want = om.step(phys.phys_id, phys.params(), q, dt, nsteps)
    return got, want, q


@pytest.mark.parametrize("bc", range(len(BCS)))
@pytest.mark.parametrize("P", [1, 2, 3])
@pytest.mark.parametrize("name", list(PHYS))
def test_fp64_parity(oracle, name, P, bc):
    print(f"fp64 {name} P={P} bc={BCS[bc]}", flush=True)
    got, want, q = run_pair(oracle, name, P, (3, 2, 2), BCS[bc], 2, M.PREC_FP64, 31 * P + bc)
    assert np.isfinite(got).all()
    assert rel_l2(got, q) > 1e-6  # the state moved
    assert rel_l2(got, want) < 1e-12


@pytest.mark.parametrize("P", [4, 5])
@pytest.mark.parametrize("name", list(PHYS))
def test_fp64_parity_high_order(oracle, name, P):
    print(f"fp64 {name} P={P}", flush=True)
    got, want, _ = run_pair(oracle, name, P, (2, 2, 2), BCS[2], 1, M.PREC_FP64, 5 + P)
    assert rel_l2(got, want) < 1e-12


@pytest.mark.parametrize("P", [1, 2, 3, 4])
@pytest.mark.parametrize("name", list(PHYS))
def test_fp32_mfma_parity(oracle, name, P):
    print(f"fp32 {name} P={P}", flush=True)
    got, want, _ = run_pair(oracle, name, P, (3, 2, 2), BCS[2], 1, M.PREC_FP32_MFMA, 7 * P)
    assert np.isfinite(got).all()
    assert rel_l2(got, want) < 1e-5


@pytest.mark.parametrize("name", list(PHYS))
def test_batch_matches_solver(name):
    print(f"batch {name}", flush=True)
    phys, P, n, L = PHYS[name], 2, (3, 2, 2), (1.0, 0.8, 1.2)
    dt = 0.1 / 3 / 5 / 3.0
    with M.Solver(n, L, phys, P=P) as s:
        q = state(name, s.num_elements, s.NP, np.random.default_rng(3))
        s.set_state(q)
        s.step_ssprk3(dt, 3)
        want = s.get_state()
    with Batch(M.PREC_FP64) as b:
        i = b.add(n, L, phys, P, q, dt, 3)
        b.step(3)
        got = b.get_state(i)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("name", list(PHYS))
def test_diagnostics_match_oracle(oracle, name):
    phys, P, n, L = PHYS[name], 3, (2, 2, 2), (1.0, 0.8, 1.2)
    with M.Solver(n, L, phys, P=P) as s:
        q = state(name, s.num_elements, s.NP, np.random.default_rng(9))
        s.set_state(q)
        got = s.diagnostics().flat()
    want = oracle.Mesh(3, P, n, L).diagnostics(phys.phys_id, q)
    nc = phys.NUM_COMPONENTS
    assert np.allclose(got[:2 * nc], want[:2 * nc], rtol=1e-12, atol=1e-14)
    assert np.array_equal(got[2 * nc:3 * nc], want[2 * nc:3 * nc])
    if name in ("sw", "mhd"):
        assert np.allclose(got[3 * nc:], want[3 * nc:], rtol=1e-12)
    else:
        assert np.isnan(got[3 * nc:]).all() and np.isnan(want[3 * nc:]).all()


def test_wrong_dimension_rejected():
    with pytest.raises((M.MxmError, ValueError)):
        M.Solver((4, 4), (1.0, 1.0), PHYS["mhd"], P=2)


@pytest.mark.parametrize("name", ["mhd", "twofluid"])
def test_limiter_parity(oracle, name):
    """The 3D limiter (all components) on the wide states, against the
    oracle's, with a state rough enough that it acts."""
    print(f"limiter {name}", flush=True)
    phys, P, n, L = PHYS[name], 2, (3, 2, 2), (1.0, 0.8, 1.2)
    dt = 0.1 * min(L) / max(n) / (2 * P + 1) / 3.0
    with M.Solver(n, L, phys, P=P) as s:
        q = state(name, s.num_elements, s.NP, np.random.default_rng(13))
        s.set_state(q)
        s.enable_cell_limiter(True, 0.05)
        s.step_ssprk3(dt, 2)
        got = s.get_state()
    want = oracle.Mesh(3, P, n, L).step(phys.phys_id, phys.params(), q, dt, 2, limiter=True, eps=0.05)
    nolim = oracle.Mesh(3, P, n, L).step(phys.phys_id, phys.params(), q, dt, 2)
    assert rel_l2(want, nolim) > 1e-3  # the limiter acted
    assert rel_l2(got, want) < 1e-12
