"""The benchmark configurations beyond C1/C2 on the GPU, through the C ABI:

* Taylor-Green (3D Euler, Rusanov, P = 3; examples/euler_taylor_green.mojo:
  86-129): 8^3 x 6 tets, 10 steps, FP64 against the oracle at 1e-12; the
  full 32^3 x 6, 200-step run in both forms, FP32-MFMA within 1e-5 of FP64
  and its kinetic energy and enstrophy (compensated FP64 sums) within 1e-7
  and 1e-6.
* The seeded Alfven-wave ensemble (2D GLM-MHD, P = 3, NC = 7;
  examples/mhd_alfven_glm_2d_gpu.mojo:98-121, 215-242): 8 of the 512
  instances, evenly spaced, in one FP64 batch for 12 steps against the
  oracle at 1e-12; all 512 in one FP32-MFMA batch to T = 1, each against
  the linear wave it approximates.
* The serving stream (dam break, Sod, PEC cavity; examples/
  shallow_water_dam_break_2d_gpu.mojo, euler_sod_2d_gpu.mojo:185-215,
  maxwell_cavity_2d_gpu.mojo:79-90): the smallest requests of each kind
  for 10 steps against the oracle at 1e-12 (limiter included), and a
  stream at the stated sizes (n in [16, 256], P in {2, 3, 4}) with end
  times cut to [0.01, 0.05] so that it runs in seconds, served with
  continuous batching and one at a time: every result equal bit for bit.
"""
import math

import numpy as np
import pytest

import mojoxm_amd as M
from mojoxm_amd import problems as pr
from mojoxm_amd.batch import Batch
from mojoxm_amd.serve import initial_state, one_at_a_time, serve

from test_gpu_parity import rel_l2

pytestmark = pytest.mark.gpu


# ---------------------------------------------------------------- C3 --------
def test_taylor_green_n8_vs_oracle(oracle):
    pb = pr.taylor_green_3d(N=8, P=3)
    with M.Solver(pb.n, pb.L, pb.physics, P=pb.P) as s:
        mesh = M.LocalMesh(*pb.n, *pb.L, P=pb.P)
        q0 = pb.ic(mesh.elem_node_xyz)
        s.set_state(q0)
        s.step_ssprk3(pb.dt, 10)
        got = s.get_state()
    want = oracle.Mesh(3, pb.P, pb.n, pb.L).step(pb.physics.phys_id, pb.physics.params(), q0, pb.dt, 10)
    assert rel_l2(got, want) <= 1e-12


def test_taylor_green_full_fp32_vs_fp64():
    pb = pr.taylor_green_3d(N=32, P=3, nsteps=200)
    mesh = M.LocalMesh(*pb.n, *pb.L, P=pb.P)
    q0 = pb.ic(mesh.elem_node_xyz).astype(np.float32).astype(np.float64)
    out, diag = {}, {}
    for prec in (M.PREC_FP64, M.PREC_FP32_MFMA):
        with M.Solver(pb.n, pb.L, pb.physics, P=pb.P, precision=prec) as s:
            s.set_state(q0)
            s.step_ssprk3(pb.dt, pb.nsteps)
            out[prec] = s.get_state()
            diag[prec] = s.diagnostics()
    assert rel_l2(out[M.PREC_FP32_MFMA], out[M.PREC_FP64]) <= 1e-5
    ke = [diag[p].kinetic_energy for p in (M.PREC_FP64, M.PREC_FP32_MFMA)]
    en = [diag[p].enstrophy for p in (M.PREC_FP64, M.PREC_FP32_MFMA)]
    assert abs(ke[1] - ke[0]) <= 1e-7 * abs(ke[0])
    assert abs(en[1] - en[0]) <= 1e-6 * abs(en[0])
    # KE of the initial field is pi^3 (exact for N >= 4); it stays within
    # the early-time decay of the inviscid run
    assert abs(ke[0] / math.pi ** 3 - 1.0) < 1e-3


# ---------------------------------------------------------------- C4 --------
def test_alfven_ensemble_instances_vs_oracle(oracle):
    members = pr.alfven_ensemble(512, seed=0)
    pick = members[::64]
    nsteps = 12
    states = {}
    with Batch(M.PREC_FP64) as b:
        ids = {}
        for m in pick:
            mesh = M.LocalMesh2D(*m.n, *m.L, P=m.P)
            q0 = m.ic(mesh.elem_node_xyz)
            states[m.index] = q0
            ids[m.index] = b.add(m.n, m.L, m.physics, m.P, q0, m.dt, nsteps)
        b.step(nsteps)
        got = {k: b.get_state(i) for k, i in ids.items()}
    for m in pick:
        want = oracle.Mesh(2, m.P, m.n, m.L).step(m.physics.phys_id, m.physics.params(), states[m.index], m.dt,
                                                  nsteps)
        assert rel_l2(got[m.index], want) <= 1e-12, m.index


def test_alfven_ensemble_full_batch_vs_linear_wave():
    """Every instance of the 512, run to T = 1 in one FP32-MFMA batch, is
    within a few per cent (relative L2 over the state) of the linear wave,
    and closer to it than the unmoved initial state wherever the wave has
    moved measurably."""
    members = pr.alfven_ensemble(512, seed=0)
    xy, ids = {}, {}
    with Batch(M.PREC_FP32_MFMA) as b:
        for m in members:
            mesh = M.LocalMesh2D(*m.n, *m.L, P=m.P)
            xy[m.index] = mesh.elem_node_xyz
            ids[m.index] = b.add(m.n, m.L, m.physics, m.P, m.ic(mesh.elem_node_xyz), m.dt, m.nsteps)
        b.step(max(m.nsteps for m in members))
        assert set(b.poll()) == set(ids.values())
        err, still = [], []
        for m in members:
            q = b.get_state(ids[m.index])
            exact = m.exact(xy[m.index], m.dt * m.nsteps)
            err.append(rel_l2(q, exact))
            still.append(rel_l2(m.ic(xy[m.index]), exact))
    err, still = np.array(err), np.array(still)
    assert np.all(np.isfinite(err))
    assert err.max() < 0.05 and err.mean() < 0.015, (err.max(), err.mean())
    moved = still > 0.01
    assert moved.sum() > 100 and np.all(err[moved] < still[moved])


# ---------------------------------------------------------------- C5 --------
def test_serving_requests_vs_oracle(oracle):
    reqs = pr.serving_stream(count=60, seed=0)
    pick = []
    for kind in ("dam_break", "sod", "cavity"):
        of = sorted((r for r in reqs if r.kind == kind), key=lambda r: r.n[0])
        pick += of[:2]
    nsteps = 10
    states = {}
    with Batch(M.PREC_FP64) as b:
        ids = {}
        for r in pick:
            q0 = initial_state(r)
            states[r.index] = q0
            ids[r.index] = b.add(r.n, r.L, r.physics, r.P, q0, r.dt, nsteps, bcs=r.bcs, limiter=r.limiter, eps=r.eps)
        b.step(nsteps)
        got = {k: b.get_state(i) for k, i in ids.items()}
    for r in pick:
        m = oracle.Mesh(2, r.P, r.n, r.L, bc=r.bcs.as_list())
        want = m.step(r.physics.phys_id, r.physics.params(), states[r.index], r.dt, nsteps, limiter=r.limiter,
                      eps=r.eps)
        assert rel_l2(got[r.index], want) <= 1e-12, (r.kind, r.n, r.P)


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_FP32_MFMA])
def test_serving_stream_alone_equals_mix(precision):
    reqs = pr.serving_stream(count=12, rate=50.0, seed=1, t_range=(0.01, 0.05))
    assert max(r.n[0] for r in reqs) > 128 and {r.P for r in reqs} == {2, 3, 4}
    states = {r.index: initial_state(r) for r in reqs}
    rep = serve(reqs, precision=precision, states=states, keep_results=True)
    _, alone = one_at_a_time(reqs, precision=precision, states=states, keep_results=True)
    assert rep.solves == len(reqs) and rep.peak_bytes <= rep.budget
    for r in reqs:
        assert np.array_equal(rep.results[r.index], alone[r.index]), (r.index, r.kind)
