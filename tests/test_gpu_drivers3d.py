"""The 3D drivers of the reference's other trait physics, run through the
C ABI on the device and checked against what the physics fixes:

  examples/maxwell_cavity.mojo      PEC standing mode, one period returns
                                    the initial state; error converges at
                                    the P = 2 rate
  examples/mhd_alfven.mojo          linear Alfven wave (small amplitude),
                                    converges at the P = 2 rate
  examples/two_fluid_langmuir.mojo  uniform plasma: the fluxes cancel and
                                    the sources give the Langmuir
                                    oscillation in closed form
  examples/shallow_water_drop.mojo  slip-walled basin: total mass is
                                    conserved

The problems are mojoxm_amd.problems.*_3d (host FP64 initial states)."""
import math

import numpy as np
import pytest

import mojoxm_amd as M
from mojoxm_amd import problems as pr

from test_gpu_parity import rel_l2

pytestmark = pytest.mark.gpu


def run(p, precision=M.PREC_FP64, q0=None):
    mesh = M.LocalMesh(*p.n, *p.L, P=p.P)
    q0 = p.ic(mesh.elem_node_xyz) if q0 is None else q0
    if precision == M.PREC_FP32_MFMA:
        q0 = q0.astype(np.float32).astype(np.float64)
    with M.Solver(p.n, p.L, p.physics, P=p.P, bcs=p.bcs, precision=precision) as s:
        s.set_state(q0)
        s.step_ssprk3(p.dt, p.nsteps)
        return mesh, q0, s.get_state()


def test_maxwell_cavity_one_period():
    """The reference driver's run: 16 x 32 x 2 cubes, P = 2, T = 2 L / c."""
    p = pr.maxwell_cavity_3d()
    print(p.name, p.nsteps, flush=True)
    mesh, q0, q = run(p)
    assert rel_l2(q, q0) < 5e-4
    assert rel_l2(q, p.exact(mesh.elem_node_xyz, p.t_final)) < 5e-4


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_FP32_MFMA])
def test_maxwell_cavity_converges(precision):
    errs = []
    for NY in (8, 16):
        p = pr.maxwell_cavity_3d(NY=NY, t_final=0.5)
        mesh, _, q = run(p, precision)
        errs.append(rel_l2(q, p.exact(mesh.elem_node_xyz, p.t_final)))
    print(errs, flush=True)
    assert errs[1] < errs[0] / 5 and errs[1] < 2e-4


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_FP32_MFMA])
def test_mhd_alfven_converges(precision):
    errs = []
    for NX in (8, 16):
        p = pr.mhd_alfven_3d(NX=NX, amp=1e-3, t_final=0.25)
        mesh, _, q = run(p, precision)
        ex = p.exact(mesh.elem_node_xyz, p.t_final)
        errs.append(rel_l2(q[..., 6], ex[..., 6]))
        # the background stays put
        assert np.abs(q[..., 5] - 1.0).max() < 1e-5
    print(errs, flush=True)
    assert errs[1] < errs[0] / 5 and errs[1] < 1e-3


def test_langmuir_oscillation_fp64():
    """Half a plasma period at the reference driver's settings (15,080
    steps): E_x, u_e and u_i against the closed form, uniform in space."""
    p = pr.langmuir_3d()
    print(p.name, p.nsteps, flush=True)
    _, _, q = run(p)
    E, ue, ui = p.notes["exact_fields"](p.t_final)
    # absolute: 1e-7 of the 0.01 velocity scale
    assert abs(q[..., 10].mean() - E) < 1e-9
    assert abs((q[..., 1] / q[..., 0]).mean() - ue) < 1e-9
    assert abs((q[..., 6] / q[..., 5]).mean() - ui) < 1e-9
    assert q[..., 10].std() < 1e-12 and np.abs(q[..., 11:13]).max() < 1e-12


def test_langmuir_oscillation_fp32_mfma():
    p = pr.langmuir_3d(t_final=math.pi / 4)
    _, _, q = run(p, M.PREC_FP32_MFMA)
    E, ue, ui = p.notes["exact_fields"](p.t_final)
    u0 = 0.01
    assert abs(q[..., 10].mean() - E) < 1e-4 * u0
    assert abs((q[..., 1] / q[..., 0]).mean() - ue) < 1e-4 * u0


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_FP32_MFMA])
def test_shallow_water_drop_conserves_mass(precision):
    p = pr.shallow_water_drop_3d()
    mesh = M.LocalMesh(*p.n, *p.L, P=p.P)
    q0 = p.ic(mesh.elem_node_xyz)
    if precision == M.PREC_FP32_MFMA:
        q0 = q0.astype(np.float32).astype(np.float64)
    with M.Solver(p.n, p.L, p.physics, P=p.P, bcs=p.bcs, precision=precision) as s:
        s.set_state(q0)
        m0 = s.diagnostics().integral[0]
        s.step_ssprk3(p.dt, p.nsteps)
        m1 = s.diagnostics().integral[0]
        q = s.get_state()
    print(m0, m1, flush=True)
    assert abs(m1 / m0 - 1.0) < (1e-12 if precision == M.PREC_FP64 else 1e-6)
    assert q[..., 0].min() > 0.9 and q[..., 0].max() < q0[..., 0].max() - 0.05


def test_frame_writer_writes_device_state(tmp_path):
    """mojoxm_amd.vtu.FrameWriter frames of a running Solver read back as
    the state's components (Float32)."""
    from mojoxm_amd import vtu
    p = pr.maxwell_cavity_3d(NY=8, t_final=0.1)
    mesh = M.LocalMesh(*p.n, *p.L, P=p.P)
    with M.Solver(p.n, p.L, p.physics, P=p.P, bcs=p.bcs) as s:
        s.set_state(p.ic(mesh.elem_node_xyz))
        fw = vtu.FrameWriter(s, mesh.elem_node_xyz, str(tmp_path), components={"Ex": 0, "Bz": 5})
        fw.write(0.0)
        s.step_ssprk3(p.dt, p.nsteps)
        path = fw.write(p.t_final)
        q = s.get_state()
        pvd = fw.finalize()
    r = vtu.read_vtu(path)
    assert np.array_equal(r["fields"]["Bz"], q[..., 5].astype(np.float32).ravel())
    assert np.abs(r["fields"]["Bz"]).max() > 0.0
    assert open(pvd).read().count("<DataSet") == 2


def test_time_loop_with_diagnostics(tmp_path):
    """run_ssprk3_loop_with_diagnostics on a running Solver: a frame and a
    CSV row at t = 0 and at each frame boundary, the last frame is the
    final state, and the state matches the same steps taken in one
    step_ssprk3 call (the clamped steps differ only by round-off)."""
    from mojoxm_amd import vtu
    p = pr.maxwell_cavity_3d(NY=8, t_final=0.2)
    mesh = M.LocalMesh(*p.n, *p.L, P=p.P)
    q0 = p.ic(mesh.elem_node_xyz)
    comps = M.DiagComponents().squared("Ex2", 0).maxabs("Bz_max", 5)
    with M.Solver(p.n, p.L, p.physics, P=p.P, bcs=p.bcs) as s:
        s.set_state(q0)
        fw = vtu.FrameWriter(s, mesh.elem_node_xyz, str(tmp_path), components={"Ex": 0})
        dw = M.DiagnosticsWriter(s, str(tmp_path / "diag.csv"), comps)
        r = M.run_ssprk3_loop_with_diagnostics(s, fw, dw, p.dt, p.t_final, num_frames=2)
        q = s.get_state()
        s.set_state(q0)
        s.step_ssprk3(p.dt, p.nsteps)
        q_ref = s.get_state()
    print(p.nsteps, r, flush=True)
    assert r.total_steps in (p.nsteps, p.nsteps + 1, p.nsteps + 2)
    assert fw.times == pytest.approx([0.0, 0.1, 0.2])
    rows = open(tmp_path / "diag.csv").read().splitlines()
    assert rows[0] == "time,Ex2,Bz_max" and len(rows) == 4
    assert float(rows[1].split(",")[2]) == 0.0 and float(rows[3].split(",")[2]) > 0.0
    last = vtu.read_vtu(str(tmp_path / fw.paths[-1]))
    assert np.array_equal(last["fields"]["Ex"], q[..., 0].astype(np.float32).ravel())
    assert rel_l2(q, q_ref) < 1e-12
