"""Parity of the HIP stage kernels (through the C ABI) with the CPU FP64
oracle on identical meshes, parameters and seeded states.

Bar: relative L2 difference <= 1e-12 after one and several SSPRK3 steps
(FP64 on both sides; the kernels sum in a different order and contract
multiply-adds, the oracle evaluates the reference's order without FMA)."""
import numpy as np
import pytest

import mojoxm_amd as M
from mojoxm_amd import physics as ph

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def random_state(phys, K, NP, rng, ndim):
    nc = phys.NUM_COMPONENTS
    if isinstance(phys, (ph.Advection, ph.Advection2D)):
        return rng.uniform(0.5, 1.5, (K, NP, 1))
    if isinstance(phys, ph.Euler):
        rho = rng.uniform(0.8, 1.2, (K, NP))
        u = rng.uniform(-0.3, 0.3, (K, NP, 3))
        if ndim == 2:
            u[..., 2] = 0.0
        p = rng.uniform(0.8, 1.2, (K, NP))
        E = p / (phys.gamma - 1) + 0.5 * rho * (u ** 2).sum(-1)
        return np.stack([rho, rho * u[..., 0], rho * u[..., 1], rho * u[..., 2], E], -1)
    if isinstance(phys, ph.Euler2D):
        rho = rng.uniform(0.8, 1.2, (K, NP))
        u = rng.uniform(-0.3, 0.3, (K, NP, 2))
        p = rng.uniform(0.8, 1.2, (K, NP))
        E = p / (phys.gamma - 1) + 0.5 * rho * (u ** 2).sum(-1)
        return np.stack([rho, rho * u[..., 0], rho * u[..., 1], E], -1)
    if isinstance(phys, ph.ShallowWater2D):
        h = rng.uniform(0.8, 1.2, (K, NP))
        u = rng.uniform(-0.3, 0.3, (K, NP, 2))
        return np.stack([h, h * u[..., 0], h * u[..., 1]], -1)
    if isinstance(phys, ph.MHDGLM2D):
        rho = rng.uniform(0.8, 1.2, (K, NP))
        u = rng.uniform(-0.2, 0.2, (K, NP, 2))
        B = rng.uniform(-0.5, 0.5, (K, NP, 2))
        p = rng.uniform(0.8, 1.2, (K, NP))
        psi = rng.uniform(-0.1, 0.1, (K, NP))
        E = p / (phys.gamma - 1) + 0.5 * rho * (u ** 2).sum(-1) + 0.5 * (B ** 2).sum(-1)
        return np.stack([rho, rho * u[..., 0], rho * u[..., 1], B[..., 0], B[..., 1], E, psi], -1)
    if isinstance(phys, ph.Maxwell2D):
        return rng.uniform(-1.0, 1.0, (K, NP, 6))
    if isinstance(phys, ph.ShallowWater):
        h = rng.uniform(0.8, 1.2, (K, NP))
        u = rng.uniform(-0.3, 0.3, (K, NP, 2))
        return np.stack([h, h * u[..., 0], h * u[..., 1]], -1)
    if isinstance(phys, ph.Maxwell):
        return rng.uniform(-1.0, 1.0, (K, NP, 6))
    if isinstance(phys, ph.IdealMHD):
        rho = rng.uniform(0.8, 1.2, (K, NP))
        u = rng.uniform(-0.2, 0.2, (K, NP, 3))
        B = rng.uniform(-0.5, 0.5, (K, NP, 3))
        p = rng.uniform(0.8, 1.2, (K, NP))
        E = p / (phys.gamma - 1) + 0.5 * rho * (u ** 2).sum(-1) + 0.5 * (B ** 2).sum(-1)
        psi = rng.uniform(-0.1, 0.1, (K, NP))
        return np.concatenate([rho[..., None], rho[..., None] * u, E[..., None], B, psi[..., None]], -1)
    if isinstance(phys, ph.FiveMomentTwoFluid):
        q = np.zeros((K, NP, 17))
        for s0, gam in ((0, phys.gamma_e), (5, phys.gamma_i)):
            rho = rng.uniform(0.8, 1.2, (K, NP))
            u = rng.uniform(-0.2, 0.2, (K, NP, 3))
            p = rng.uniform(0.8, 1.2, (K, NP))
            q[..., s0] = rho
            q[..., s0 + 1:s0 + 4] = rho[..., None] * u
            q[..., s0 + 4] = p / (gam - 1) + 0.5 * rho * (u ** 2).sum(-1)
        q[..., 10:17] = rng.uniform(-0.5, 0.5, (K, NP, 7))
        return q
    raise TypeError(phys)


def run_pair(oracle, n, L, P, phys, bc=None, nsteps=1, dt=None, seed=0, limiter=False, eps=0.1, q0=None):
    ndim = len(n)
    bcs = (M.BoundaryConditions(*bc) if ndim == 3 else M.BoundaryConditions2D(*bc)) if bc else None
    with M.Solver(n, L, phys, P=P, bcs=bcs) as s:
        rng = np.random.default_rng(seed)
        q = random_state(phys, s.num_elements, s.NP, rng, ndim) if q0 is None else q0
        if dt is None:
            dt = 0.2 * min(L) / max(n) / (2 * P + 1) / 2.0
        om = oracle.Mesh(ndim, P, n, L, bc)
        s.set_state(q)
        if limiter:
            s.enable_cell_limiter(True, eps)
        s.step_ssprk3(dt, nsteps)
        got = s.get_state()
    want = om.step(phys.phys_id, phys.params(), q, dt, nsteps, limiter=limiter, eps=eps)
    return got, want, q


CASES_3D = [
    (ph.Advection(1.0, 0.7, -0.4), None),
    (ph.Advection(1.0, 0.7, -0.4, inflow_q=0.3), [3, 2, 1, 2, 3, 1]),
    (ph.Euler(flux_type=ph.FLUX_RUSANOV), None),
    (ph.Euler(flux_type=ph.FLUX_ROE), None),
    (ph.Euler(flux_type=ph.FLUX_HLLE), None),
    (ph.Euler(flux_type=ph.FLUX_HLLEC), None),
    (ph.Euler(flux_type=ph.FLUX_ROE, entropy_fix=True), None),
    (ph.Euler(flux_type=ph.FLUX_HLLEC, entropy_fix=True, gx=0.1, gy=-0.2, gz=0.3), None),
    (ph.Euler(flux_type=ph.FLUX_RUSANOV, inflow=(1.0, 0.1, 0.0, 0.0, 2.5)), [1, 2, 3, 1, 2, 0]),
    (ph.Euler(flux_type=ph.FLUX_HLLEC, inflow=(1.0, 0.1, 0.0, 0.0, 2.5)), [0, 1, 0, 2, 3, 0]),
]


@pytest.mark.parametrize("case", range(len(CASES_3D)))
@pytest.mark.parametrize("P", [1, 2, 3])
def test_stage3d_parity(oracle, case, P):
    phys, bc = CASES_3D[case]
    got, want, _ = run_pair(oracle, (3, 2, 3), (1.0, 0.8, 1.2), P, phys, bc, nsteps=2, seed=case * 7 + P)
    assert np.isfinite(got).all()
    assert rel_l2(got, want) < TOL


@pytest.mark.parametrize("P", [4, 5])
@pytest.mark.parametrize("flux", [ph.FLUX_RUSANOV, ph.FLUX_HLLEC])
def test_stage3d_parity_high_order(oracle, P, flux):
    got, want, _ = run_pair(oracle, (2, 2, 2), (1.0, 1.0, 1.0), P, ph.Euler(flux_type=flux), None, nsteps=1)
    assert rel_l2(got, want) < TOL


CASES_2D = [
    (ph.Advection(1.0, -0.6), None),
    (ph.Advection(1.0, -0.6, inflow_q=0.4), [3, 2, 1, 3]),
    (ph.Euler(flux_type=ph.FLUX_RUSANOV), None),
    (ph.Euler(flux_type=ph.FLUX_ROE), None),
    (ph.Euler(flux_type=ph.FLUX_HLLE), None),
    (ph.Euler(flux_type=ph.FLUX_HLLEC), None),
    (ph.Euler(flux_type=ph.FLUX_HLLE, entropy_fix=True), [2, 2, 1, 1]),
    (ph.Euler(flux_type=ph.FLUX_HLLEC, inflow=(1.0, 0.2, 0.0, 0.0, 2.5)), [3, 2, 1, 1]),
    (ph.Euler2D(), None),
    (ph.Euler2D(hllc=True, gx=0.1, gy=-0.3), None),
    (ph.Euler2D(hllc=True, inflow=(1.0, 0.2, 0.0, 2.5)), [3, 2, 1, 1]),
    (ph.ShallowWater2D(), None),
    (ph.ShallowWater2D(hll=True), [1, 2, 1, 1]),
    (ph.ShallowWater2D(inflow=(1.1, 0.2, 0.0)), [3, 2, 1, 1]),
    (ph.MHDGLM2D(c_h=1.5, alpha_d=1.5 / 0.18), None),
    (ph.MHDGLM2D(c_h=1.2, inflow=(1.0, 0.1, 0.0, 0.3, 0.1, 2.0, 0.0)), [3, 2, 1, 1]),
    (ph.Maxwell2D(c=1.0, J=(0.1, -0.2, 0.05), M=(0.0, 0.1, -0.1)), None),
    (ph.Maxwell2D(c=1.0), [1, 1, 1, 1]),
    (ph.Maxwell2D(c=1.0, inflow=(0.1, 0.0, 0.2, 0.0, 0.1, 0.0)), [3, 2, 0, 0]),
    (ph.Advection2D(0.8, -1.0), None),
    (ph.Advection2D(0.8, -1.0, inflow_q=0.5), [3, 1, 2, 3]),
]


@pytest.mark.parametrize("case", range(len(CASES_2D)))
@pytest.mark.parametrize("P", [1, 2, 3, 4])
def test_stage2d_parity(oracle, case, P):
    phys, bc = CASES_2D[case]
    got, want, _ = run_pair(oracle, (4, 3), (2.0, 1.5), P, phys, bc, nsteps=2, seed=case * 11 + P)
    assert np.isfinite(got).all()
    assert rel_l2(got, want) < TOL


@pytest.mark.parametrize("case", [2, 5, 8, 11, 14, 16])
def test_stage2d_parity_p5(oracle, case):
    phys, bc = CASES_2D[case]
    got, want, _ = run_pair(oracle, (3, 3), (1.0, 1.0), 5, phys, bc, nsteps=1, seed=case)
    assert rel_l2(got, want) < TOL


@pytest.mark.parametrize("ndim", [2, 3])
def test_limiter_parity(oracle, ndim):
    """Shock-like data so that theta < 1 on many elements."""
    if ndim == 3:
        n, L, phys = (3, 3, 2), (1.0, 1.0, 1.0), ph.Euler(flux_type=ph.FLUX_HLLEC)
    else:
        n, L, phys = (6, 4), (1.0, 1.0), ph.Euler2D(hllc=True)
    rng = np.random.default_rng(5)
    with M.Solver(n, L, phys, P=2) as s:
        q = random_state(phys, s.num_elements, s.NP, rng, ndim)
    q[: q.shape[0] // 2, :, 0] *= 3.0  # density jump
    got, want, _ = run_pair(oracle, n, L, 2, phys, None, nsteps=2, limiter=True, q0=q)
    assert rel_l2(got, want) < TOL
    # the limiter did act (the result differs from the unlimited step)
    un, _, _ = run_pair(oracle, n, L, 2, phys, None, nsteps=2, limiter=False, q0=q)
    assert rel_l2(got, un) > 1e-6


@pytest.mark.parametrize("P", [2, 3, 4, 5])
def test_constant_state_preserved(P):
    """test/euler_2d_gpu_test.mojo:5-18 and test/euler_3d_test.mojo: a uniform
    state is a fixed point of every stage (to round-off)."""
    for n, L, phys, q in [
        ((3, 2, 2), (1.0, 1.0, 1.0), ph.Euler(flux_type=ph.FLUX_HLLEC), [1.0, 0.5, 0.3, -0.2, 2.5 + 0.19]),
        ((4, 4), (1.0, 1.0), ph.Euler2D(hllc=True), [1.0, 0.5, 0.3, 2.5 + 0.17]),
        ((4, 4), (1.0, 1.0), ph.ShallowWater2D(hll=True), [1.0, 0.2, -0.1]),
        ((4, 4), (1.0, 1.0), ph.MHDGLM2D(c_h=1.0), [1.0, 0.1, 0.2, 0.3, -0.1, 2.0, 0.0]),
        ((4, 4), (1.0, 1.0), ph.Maxwell2D(), [0.1, 0.2, 0.3, -0.1, 0.2, 0.4]),
    ]:
        with M.Solver(n, L, phys, P=P) as s:
            q0 = np.broadcast_to(np.array(q), s.shape).copy()
            s.set_state(q0)
            s.step_ssprk3(1e-3, 3)
            got = s.get_state()
        assert np.abs(got - q0).max() < 1e-12 * np.abs(q0).max(), type(phys).__name__


def test_config1_full_run_parity(oracle):
    """C1: 16^3 x 6 tets, P=2, upwind advection of the Gaussian, 462 steps to
    T=1; the GPU state matches the oracle's and returns to the IC."""
    pb = M.problems.advection_gaussian_3d()
    assert pb.nsteps == 462
    with M.Solver(pb.n, pb.L, pb.physics, P=pb.P) as s:
        mesh = M.LocalMesh(*pb.n, *pb.L, P=pb.P)
        q0 = pb.ic(mesh.elem_node_xyz)
        s.set_state(q0)
        s.step_ssprk3(pb.dt, pb.nsteps)
        got = s.get_state()
    om = oracle.Mesh(3, pb.P, pb.n, pb.L)
    want = om.step(pb.physics.phys_id, pb.physics.params(), q0, pb.dt, pb.nsteps)
    assert rel_l2(got, want) < TOL
    err = rel_l2(got, q0)
    assert err < 0.1  # DG P=2 on 16^3 damps the sigma=0.1 pulse by a few percent


@pytest.mark.parametrize("n", [(1, 2, 2), (2, 1, 2), (2, 2, 1), (1, 1, 1), (1, 3), (4, 1), (1, 1)])
@pytest.mark.parametrize("P", [1, 3])
def test_one_cell_per_period(oracle, n, P):
    """A periodic axis of one cell: the two sides of a face are the same
    cell's opposite faces, in the same tile."""
    L = (1.0, 0.8, 1.2)[: len(n)]
    phys = ph.Euler(flux_type=ph.FLUX_HLLEC)
    bc = [0, 0, 1, 1, 0, 0][: 2 * len(n)]
    got, want, _ = run_pair(oracle, n, L, P, phys, bc, nsteps=2, seed=P + len(n))
    assert rel_l2(got, want) < TOL
