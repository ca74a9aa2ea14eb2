"""The Solver API that a host binding calls besides step_ssprk3
(src/solver.mojo:711-729, 757-914, 961-1005, 1071-1092), through the C ABI:

* three rk_stage calls equal one step_ssprk3 bit for bit (FP64 and FP32-MFMA,
  with and without the limiter, 2D and 3D, GLM damping at stage 3);
* memory_report: every category but the halo positive, the halo zero, the
  total equal to the category sum (test/memory_report_test.mojo:59-76), the
  stage trio 3 * slots * NP * NC * word with slots >= K, and the device
  memory the solver holds (hipMemGetInfo) at least the reported total;
* dof_count = K * NP * NC and state_bytes_per_step = 8 * dof * word
  (src/solver.mojo:757-763, 804-824);
* download_owned_component(c) equals component c of the state;
* bench_step_loop reports its own arithmetic and advances the state by
  warmup + measure steps;
* time_stage_kernel leaves the limiter setting as it found it.
"""
import ctypes as C

import numpy as np
import pytest

import mojoxm_amd as M
from mojoxm_amd import physics as ph

from test_gpu_parity import random_state

pytestmark = pytest.mark.gpu

CASES = [
    ("euler2d_hllec_walls", (6, 5), (1.0, 0.8), lambda: ph.Euler(flux_type=ph.FLUX_HLLEC), 3,
     M.BoundaryConditions2D(2, 2, 1, 1)),
    ("glm2d_damped", (5, 4), (1.0, 1.0), lambda: ph.MHDGLM2D(c_h=1.3, alpha_d=1.3 / 0.18), 3, None),
    ("euler3d_roe", (3, 2, 2), (1.0, 1.0, 1.0), lambda: ph.Euler(flux_type=ph.FLUX_ROE, entropy_fix=True), 2, None),
    ("sw2d_hll", (4, 6), (1.0, 1.0), lambda: ph.ShallowWater2D(hll=True), 2, M.BoundaryConditions2D(1, 1, 1, 1)),
]


def make(case, precision, rng):
    name, n, L, phys_f, P, bcs = case
    phys = phys_f()
    s = M.Solver(n, L, phys, P=P, bcs=bcs, precision=precision)
    q0 = random_state(phys, s.num_elements, s.NP, rng, len(n))
    if precision == M.PREC_FP32_MFMA:
        q0 = q0.astype(np.float32).astype(np.float64)
    return s, q0


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_FP32_MFMA])
@pytest.mark.parametrize("limiter", [False, True])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_rk_stage_three_times_equals_step(case, precision, limiter):
    rng = np.random.default_rng(7)
    dt = 1e-3
    a, q0 = make(case, precision, rng)
    b, _ = make(case, precision, rng)
    with a, b:
        for s in (a, b):
            s.set_state(q0)
            if limiter:
                s.enable_cell_limiter(True, 0.1)
        for _ in range(3):
            for st in range(3):
                a.rk_stage(st, dt)
        b.step_ssprk3(dt, 3)
        qa, qb = a.get_state(), b.get_state()
    assert np.array_equal(qa, qb)
    assert not np.array_equal(qa, q0)


def _mem_free():
    lib = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert lib.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_FP32_MFMA])
@pytest.mark.parametrize("dim,P", [(3, 2), (3, 3), (3, 4), (3, 5), (2, 3)])
def test_memory_report_and_dof(dim, P, precision):
    word = 8 if precision == M.PREC_FP64 else 4
    n, L = ((4, 4, 4), (1.0,) * 3) if dim == 3 else ((24, 24), (1.0, 1.0))
    phys = ph.Advection(1.0, 0.0, 0.0) if dim == 3 else ph.Euler()
    free0 = _mem_free()
    with M.Solver(n, L, phys, P=P, precision=precision) as s:
        held = free0 - _mem_free()
        rep = s.memory_report()
        K, NP, NC = s.num_elements, s.NP, s.NC
        assert rep.rk_stage_bytes > 0 and rep.dg_operators_bytes > 0
        assert rep.limiter_bytes > 0 and rep.mesh_connectivity_bytes > 0
        assert rep.halo_device_bytes == 0 and rep.halo_pinned_bytes == 0
        assert rep.total_device_bytes() == (rep.rk_stage_bytes + rep.dg_operators_bytes + rep.limiter_bytes +
                                            rep.mesh_connectivity_bytes + rep.halo_device_bytes)
        slots, rem = divmod(rep.rk_stage_bytes, 3 * NP * NC * word)
        assert rem == 0 and K <= slots <= 2 * K
        assert held >= rep.total_device_bytes()
        assert s.dof_count() == K * NP * NC
        assert s.state_bytes_per_step() == 8 * K * NP * NC * word
        rep.print()


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_FP32_MFMA])
def test_download_component(precision):
    rng = np.random.default_rng(11)
    s, q0 = make(CASES[1], precision, rng)
    with s:
        s.set_state(q0)
        s.step_ssprk3(1e-3, 2)
        q = s.get_state()
        for c in range(s.NC):
            assert np.array_equal(s.download_owned_component(c), q[:, :, c].ravel())


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_FP32_MFMA])
def test_bench_step_loop(precision):
    rng = np.random.default_rng(13)
    dt = 1e-3
    a, q0 = make(CASES[0], precision, rng)
    b, _ = make(CASES[0], precision, rng)
    with a, b:
        a.set_state(q0)
        b.set_state(q0)
        t = a.bench_step_loop(dt, warmup_steps=2, measure_steps=10)
        b.step_ssprk3(dt, 12)
        assert np.array_equal(a.get_state(), b.get_state())
        assert t.steps == 10 and t.wall_seconds > 0
        assert t.dof == a.dof_count() and t.state_bytes_per_step == a.state_bytes_per_step()
        assert t.dof_per_second() == pytest.approx(t.dof * 10 / t.wall_seconds, rel=1e-12)
        assert t.dof_stage_updates_per_second() == pytest.approx(3 * t.dof_per_second(), rel=1e-12)
        # device-side figures: the events sit inside the wall-clock interval
        assert 0 < 30 * t.stage_seconds <= t.wall_seconds * (1 + 1e-6)
        assert t.achieved_hbm_gbs == pytest.approx(t.state_bytes_per_step / 3 / t.stage_seconds * 1e-9, rel=1e-12)
        assert t.hbm_fraction == pytest.approx(t.achieved_hbm_gbs / 8000.0, rel=1e-12)
        if precision == M.PREC_FP32_MFMA:
            nd, npn = 2, a.NP  # CASES[0] is 2D P=3: 3 faces of P+1 nodes
            flops = 2.0 * t.dof * (nd * npn + 3 * 4)
            assert t.mfma_tflops == pytest.approx(flops / t.stage_seconds * 1e-12, rel=1e-12)
            assert t.mfma_utilization == pytest.approx(t.mfma_tflops / 157.3, rel=1e-12)
        else:
            assert t.mfma_tflops == 0.0 and t.mfma_utilization == 0.0
        t.print()


def test_time_stage_kernel_restores_limiter():
    rng = np.random.default_rng(17)
    dt = 1e-3
    a, q0 = make(CASES[0], M.PREC_FP64, rng)
    b, _ = make(CASES[0], M.PREC_FP64, rng)
    with a, b:
        for s in (a, b):
            s.set_state(q0)
            s.enable_cell_limiter(True, 0.1)
        a.time_stage_kernel(dt, 2)
        a.set_state(q0)
        a.step_ssprk3(dt, 2)
        b.step_ssprk3(dt, 2)
        assert np.array_equal(a.get_state(), b.get_state())


@pytest.mark.parametrize("limiter", [False, True])
def test_prepare_steps_leaves_the_state_and_the_steps_unchanged(limiter):
    # prepare_steps builds graphs without stepping: the state is unchanged,
    # and the steps that follow give the bits of a solver that never
    # prepared
    rng = np.random.default_rng(23)
    dt = 1e-3
    a, q0 = make(CASES[0], M.PREC_FP64, rng)
    b, _ = make(CASES[0], M.PREC_FP64, rng)
    with a, b:
        for s in (a, b):
            s.set_state(q0)
            if limiter:
                s.enable_cell_limiter(True, 0.1)
        a.prepare_steps(dt, 37)
        assert np.array_equal(a.get_state(), q0)
        a.step_ssprk3(dt, 37)
        b.step_ssprk3(dt, 37)
        assert np.array_equal(a.get_state(), b.get_state())
