"""Device diagnostics (mxm_dg_diagnostics, mxm_batch_diagnostics) against
the oracle's (oracle/diag.c, long-double sums) on the same seeded states.

Both evaluate the same mass-weighted nodal rule in (at least) FP64; the
device sums per node in compensated FP64 in a fixed order, so the stated
bar is a relative difference <= 1e-13 per value (atol 1e-13 x the value's
scale for sums that cancel to ~0), identical bits on repeated calls, and
identical bits for an instance of a batch and the same state in a Solver.
The FP32-MFMA form is checked on its FP32 state: the diagnostics are
still computed in FP64 from it."""
import math

import numpy as np
import pytest

import mojoxm_amd as M
from mojoxm_amd import physics as ph
from mojoxm_amd import problems
from mojoxm_amd.batch import Batch

from test_gpu_parity import random_state

pytestmark = pytest.mark.gpu

TOL = 1e-13

CASES = [
    ((5, 4), (2.0, 3.0), ph.Euler(flux_type=ph.FLUX_HLLEC), 3),
    ((4, 4), (1.0, 1.0), ph.Euler2D(), 2),
    ((4, 5), (1.0, 1.0), ph.MHDGLM2D(c_h=1.0, alpha_d=1.0), 3),
    ((6, 3), (1.0, 2.0), ph.ShallowWater2D(), 4),
    ((4, 4), (1.0, 1.0), ph.Maxwell2D(), 2),
    ((3, 2, 2), (1.0, 1.0, 2.0), ph.Euler(flux_type=ph.FLUX_RUSANOV), 3),
    ((2, 2, 3), (1.0, 1.0, 1.0), ph.Advection(1.0, 0.5, 0.2), 2),
    ((2, 2, 2), (1.0, 1.0, 1.0), ph.Euler(), 5),
]


def check_close(got, want, nc):
    got, want = np.asarray(got), np.asarray(want)
    if np.isnan(want[3 * nc]):
        assert np.isnan(got[3 * nc]) and np.isnan(got[3 * nc + 1])
        got, want = got[:3 * nc], want[:3 * nc]
    # scale of each sum: the matching square integral bounds |int q_c|
    scale = np.abs(want).copy()
    scale[:nc] = np.maximum(scale[:nc], np.sqrt(np.abs(want[nc:2 * nc])))
    assert np.all(np.abs(got - want) <= TOL * np.maximum(scale, 1e-300) + TOL * 1e-3), \
        np.max(np.abs(got - want) / np.maximum(scale, 1e-300))


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_FP32_MFMA])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_diagnostics_vs_oracle(oracle, case, precision):
    n, L, phys, P = CASES[case]
    ndim = len(n)
    nc = phys.NUM_COMPONENTS
    with M.Solver(n, L, phys, P=P, precision=precision) as s:
        q = random_state(phys, s.num_elements, s.NP, np.random.default_rng(40 + case), ndim)
        if precision == M.PREC_FP32_MFMA:
            q = q.astype(np.float32).astype(np.float64)
        s.set_state(q)
        d1 = s.diagnostics().flat()
        d2 = s.diagnostics().flat()
    assert np.array_equal(d1, d2, equal_nan=True)
    want = oracle.Mesh(ndim, P, n, L).diagnostics(phys.phys_id, q)
    check_close(d1, want, nc)


def test_batch_instance_diagnostics_equal_solver():
    phys = ph.Euler(flux_type=ph.FLUX_HLLEC)
    rng = np.random.default_rng(7)
    qs = {nn: random_state(phys, 2 * nn * nn, 10, rng, 2) for nn in (6, 9)}
    with Batch(M.PREC_FP64) as b:
        ids = {nn: b.add((nn, nn), (1.0, 1.0), phys, 3, qs[nn], 1e-3, 2) for nn in qs}
        got = {nn: b.diagnostics(i).flat() for nn, i in ids.items()}
    for nn, q in qs.items():
        with M.Solver((nn, nn), (1.0, 1.0), phys, P=3) as s:
            s.set_state(q)
            assert np.array_equal(s.diagnostics().flat(), got[nn])


def test_taylor_green_initial_ke_enstrophy():
    """C3's initial field on a coarse mesh: KE = pi^3 exactly for the nodal
    rule, enstrophy near 3 pi^3 (interpolation error of the derivative)."""
    pb = problems.taylor_green_3d(N=8)
    mesh = M.LocalMesh(*pb.n, *pb.L, P=pb.P)
    with M.Solver(pb.n, pb.L, pb.physics, P=pb.P) as s:
        s.set_state(pb.ic(mesh.elem_node_xyz))
        d = s.diagnostics()
    assert d.kinetic_energy == pytest.approx(math.pi ** 3, rel=1e-12)
    assert d.enstrophy == pytest.approx(3 * math.pi ** 3, rel=1e-2)
    assert d.integral[0] == pytest.approx((2 * math.pi) ** 3, rel=1e-12)
