"""The FP32-MFMA form (MXM_PREC_FP32_MFMA) against the CPU FP64 oracle.

FP32 state, traces and Riemann fluxes with the volume and LIFT
contractions on V_MFMA_F32_16X16X4F32 (FP32 accumulation); geometry, the
RK combine and the positivity rescue in FP64.  The stated bar is a
relative L2 difference <= 1e-5 from the FP64 oracle on the same mesh and
seeded state (north-star accuracy target of the FP32 form); the observed
differences are at the FP32 rounding level, ~1e-7.  Every P exercises the
MFMA operand layout, including the multi-tile rows of NP > 16 (3D P >= 3,
2D P = 5)."""
import numpy as np
import pytest

import mojoxm_amd as M
from mojoxm_amd import physics as ph

from test_gpu_parity import CASES_2D, CASES_3D, random_state, rel_l2

pytestmark = pytest.mark.gpu

TOL32 = 1e-5


def run_pair32(oracle, n, L, P, phys, bc=None, nsteps=2, seed=0, limiter=False, eps=0.1):
    ndim = len(n)
    bcs = (M.BoundaryConditions(*bc) if ndim == 3 else M.BoundaryConditions2D(*bc)) if bc else None
    with M.Solver(n, L, phys, P=P, bcs=bcs, precision=M.PREC_FP32_MFMA) as s:
        rng = np.random.default_rng(seed)
        q = random_state(phys, s.num_elements, s.NP, rng, ndim)
        # the oracle starts from the same FP32-representable state
        q = q.astype(np.float32).astype(np.float64)
        dt = 0.2 * min(L) / max(n) / (2 * P + 1) / 2.0
        s.set_state(q)
        if limiter:
            s.enable_cell_limiter(True, eps)
        s.step_ssprk3(dt, nsteps)
        got = s.get_state()
    om = oracle.Mesh(ndim, P, n, L, bc)
    want = om.step(phys.phys_id, phys.params(), q, dt, nsteps, limiter=limiter, eps=eps)
    return got, want


@pytest.mark.parametrize("case", range(len(CASES_3D)))
@pytest.mark.parametrize("P", [1, 2, 3])
def test_fp32_stage3d(oracle, case, P):
    phys, bc = CASES_3D[case]
    got, want = run_pair32(oracle, (3, 2, 3), (1.0, 0.8, 1.2), P, phys, bc, seed=case * 7 + P)
    assert np.isfinite(got).all()
    assert rel_l2(got, want) < TOL32


@pytest.mark.parametrize("P", [4, 5])
@pytest.mark.parametrize("flux", [ph.FLUX_RUSANOV, ph.FLUX_HLLEC])
def test_fp32_stage3d_high_order(oracle, P, flux):
    got, want = run_pair32(oracle, (2, 2, 2), (1.0, 1.0, 1.0), P, ph.Euler(flux_type=flux))
    assert rel_l2(got, want) < TOL32


@pytest.mark.parametrize("case", range(len(CASES_2D)))
@pytest.mark.parametrize("P", [1, 2, 3, 4, 5])
def test_fp32_stage2d(oracle, case, P):
    phys, bc = CASES_2D[case]
    got, want = run_pair32(oracle, (4, 3), (2.0, 1.5), P, phys, bc, seed=case * 11 + P)
    assert np.isfinite(got).all()
    assert rel_l2(got, want) < TOL32


@pytest.mark.parametrize("ndim", [2, 3])
def test_fp32_limiter(oracle, ndim):
    """Limiter after every stage on a smooth random state: the means and
    theta are FP64 on the GPU too."""
    if ndim == 2:
        got, want = run_pair32(oracle, (6, 4), (1.0, 1.0), 2, ph.Euler(flux_type=ph.FLUX_HLLE), None,
                               nsteps=3, limiter=True)
    else:
        got, want = run_pair32(oracle, (3, 3, 2), (1.0, 1.0, 1.0), 2, ph.Euler(flux_type=ph.FLUX_HLLEC), None,
                               nsteps=2, limiter=True)
    assert rel_l2(got, want) < TOL32


def test_fp32_vortex_run(oracle):
    """64 steps of the isentropic vortex (32^2 squares, P = 3, HLLEC):
    error against the FP64 oracle stays at the FP32 level and the solution
    error against the analytic vortex matches the FP64 form's."""
    pb = M.problems.euler_vortex_2d(N=32, t_final=64 * M.problems.euler_vortex_2d(N=32).dt)
    mesh = M.LocalMesh2D(*pb.n, *pb.L, P=pb.P)
    q0 = pb.ic(mesh.elem_node_xyz).astype(np.float32).astype(np.float64)
    out = {}
    for prec in (M.PREC_FP64, M.PREC_FP32_MFMA):
        with M.Solver(pb.n, pb.L, pb.physics, P=pb.P, precision=prec) as s:
            s.set_state(q0)
            s.step_ssprk3(pb.dt, pb.nsteps)
            out[prec] = s.get_state()
    om = oracle.Mesh(2, pb.P, pb.n, pb.L)
    want = om.step(pb.physics.phys_id, pb.physics.params(), q0, pb.dt, pb.nsteps)
    assert rel_l2(out[M.PREC_FP64], want) < 1e-12
    assert rel_l2(out[M.PREC_FP32_MFMA], want) < TOL32
    exact = pb.exact(mesh.elem_node_xyz, pb.t_final)
    e64, e32 = rel_l2(out[M.PREC_FP64], exact), rel_l2(out[M.PREC_FP32_MFMA], exact)
    assert abs(e32 - e64) < 0.01 * e64


def test_fp32_error_budget_full_vortex_run():
    """The error bar of the FP32-MFMA form on its benchmark run: the
    128^2 P = 3 vortex to T = 2 (4,054 steps) from the same
    FP32-representable state stays within 1e-5 relative L2 of the plain FP64
    kernels (observed 5.2e-6 with the stochastically rounded state store
    and the upwind-side HLLEC flux, 3.0e-6 with the wave-sum flux, 9.67e-6
    with round-to-nearest storage)."""
    pb = M.problems.euler_vortex_2d()
    mesh = M.LocalMesh2D(*pb.n, *pb.L, P=pb.P)
    q0 = pb.ic(mesh.elem_node_xyz).astype(np.float32).astype(np.float64)
    out = {}
    for prec in (M.PREC_FP64, M.PREC_FP32_MFMA):
        print(f"vortex T=2 precision {prec}", flush=True)
        with M.Solver(pb.n, pb.L, pb.physics, P=pb.P, precision=prec) as s:
            s.set_state(q0)
            s.step_ssprk3(pb.dt, pb.nsteps)
            out[prec] = s.get_state()
    assert rel_l2(out[M.PREC_FP32_MFMA], out[M.PREC_FP64]) < 1e-5


def test_fp32_error_budget_largest_alfven_instances():
    """The same bar on the longest runs of the 512-instance GLM-MHD Alfven
    ensemble (n = 95-96, 4,991-5,529 steps to T = 1, more than the vortex's
    4,054) and on two shorter ones: each FP32-MFMA final state within 1e-5
    relative L2 of the plain FP64 one from the same FP32-representable
    state.  With round-to-nearest storage the n = 96 instance drifted to
    1.19e-5; the stochastic rounding of the stored state keeps every one
    below 7e-6."""
    members = M.problems.alfven_ensemble(512, seed=0)
    pick = sorted(members, key=lambda m: (-m.n[0], m.index))[:3] + [members[0], members[384]]
    for m in pick:
        mesh = M.LocalMesh2D(*m.n, *m.L, P=m.P)
        q0 = m.ic(mesh.elem_node_xyz).astype(np.float32).astype(np.float64)
        out = {}
        for prec in (M.PREC_FP64, M.PREC_FP32_MFMA):
            print(f"alfven instance {m.index} n={m.n[0]} steps={m.nsteps} precision {prec}", flush=True)
            with M.Solver(m.n, m.L, m.physics, P=m.P, precision=prec) as s:
                s.set_state(q0)
                s.step_ssprk3(m.dt, m.nsteps)
                out[prec] = s.get_state()
        assert rel_l2(out[M.PREC_FP32_MFMA], out[M.PREC_FP64]) < TOL32, m.index


def test_fp32_stored_state_rounding_is_unbiased():
    """The stochastic rounding of the FP32 form's stored state keeps the
    expected value.  A uniform 2D Euler state (rho = 1, u = (1, 0), p = 1)
    under a uniform gravity g along x: the fluxes stay uniform, so
    m_x(t) = 1 + g t exactly.  With dt g a quarter of a float ulp of m_x,
    round-to-nearest storage (or truncation) would keep m_x at exactly 1,
    and a draw biased upward would add an ulp per store; the mean of m_x
    over the nodes follows the FP64 run (measured 1.6 % above it, for a
    500-ulp rise).  The 10 % bound covers the FP32 operator's own drift on
    the uniform state (its volume and surface sums cancel only to FP32
    rounding: -7.8e-7 in m_x over the same run with g = 0, 1.3 % of the
    rise) and the random walk of ten node classes (the salt depends on the
    node only)."""
    g, dt, nsteps = 3.0e-5, 1.0e-3, 2000
    phys = ph.Euler(gx=g)
    n = (8, 8)
    rise = {}
    for prec in (M.PREC_FP64, M.PREC_FP32_MFMA):
        with M.Solver(n, (1.0, 1.0), phys, P=3, precision=prec) as s:
            q = np.zeros(s.shape)
            q[..., 0], q[..., 1], q[..., 4] = 1.0, 1.0, 1.0 / 0.4 + 0.5
            s.set_state(q)
            s.step_ssprk3(dt, nsteps)
            rise[prec] = s.get_state()[..., 1].mean() - 1.0
    assert rise[M.PREC_FP64] == pytest.approx(g * dt * nsteps, rel=1e-6)
    assert rise[M.PREC_FP32_MFMA] == pytest.approx(rise[M.PREC_FP64], rel=0.1)
