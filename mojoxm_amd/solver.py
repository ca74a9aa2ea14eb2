"""Device-resident DG solver: the Python face of Solver[PhysT, P]
(src/solver.mojo:538-1093) and of the 2D <phys>_rk_stage_2d drivers, over
the C ABI of libmxm_hip.so.  State lives on the GPU; the host sees FP64
arrays shaped [K, NP, NC] (element-major, src/solver.mojo:17-21)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .boundary import BoundaryConditions, BoundaryConditions2D
from .memory_report import MemoryReport, ThroughputReport

PREC_FP64 = 0       # plain FP64 (include/mxm_dg.h MXM_PREC_FP64)
PREC_FP32_MFMA = 1  # FP32 state and fluxes, volume + LIFT on FP32 MFMA (MXM_PREC_FP32_MFMA)


def launch_config_count(precision: int) -> int:
    """Launch configurations of the stage kernels of one arithmetic form."""
    v = C.c_int()
    L.check(L.lib().mxm_get_constant(2 if precision == PREC_FP64 else 3, C.byref(v)), "mxm_get_constant")
    return v.value


class Solver:
    """DG solver on a periodic Kuhn-tet (ndim=3) or /-split triangle
    (ndim=2) mesh.

    Parameters mirror the reference's Driver3D / 2D driver composition:
    mesh size and extent, boundary conditions, order P and a physics
    descriptor from ``mojoxm_amd.physics``.  ``precision`` selects the
    plain FP64 kernels (PREC_FP64) or the FP32-MFMA form (PREC_FP32_MFMA);
    the host side always sees FP64 arrays.
    """

    def __init__(self, n, L_box, physics, P: int = 2, bcs=None, precision: int = PREC_FP64, device: int = 0):
        L.ensure_device(device)
        self.ndim = len(n)
        if self.ndim not in physics.dims:
            raise ValueError(f"{type(physics).__name__} does not run on a {self.ndim}D mesh")
        self.physics, self.P, self.n, self.L = physics, P, tuple(n), tuple(L_box)
        self.precision = precision
        self._prm = np.ascontiguousarray(physics.params(), dtype=np.float64)
        h = C.c_void_p()
        if self.ndim == 3:
            bcs = bcs or BoundaryConditions.periodic()
            d = L.Mesh3DDesc(n[0], n[1], n[2], L_box[0], L_box[1], L_box[2], (C.c_int * 6)(*bcs.as_list()))
            L.check(L.lib().mxm_dg_create_3d(C.byref(d), P, physics.phys_id, self._prm.ctypes.data, precision,
                                             C.byref(h)), "mxm_dg_create_3d")
        else:
            bcs = bcs or BoundaryConditions2D.periodic()
            d = L.Mesh2DDesc(n[0], n[1], L_box[0], L_box[1], (C.c_int * 4)(*bcs.as_list()))
            L.check(L.lib().mxm_dg_create_2d(C.byref(d), P, physics.phys_id, self._prm.ctypes.data, precision,
                                             C.byref(h)), "mxm_dg_create_2d")
        self.bcs = bcs
        self._h = h
        info = (C.c_int * 5)()
        L.check(L.lib().mxm_dg_info(h, info), "mxm_dg_info")
        _, self.num_elements, self.NP, self.NC, self.num_faces = list(info)

    # ---- lifetime ----------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            L.lib().mxm_dg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- state ---------------------------------------------------------------
    @property
    def shape(self):
        return (self.num_elements, self.NP, self.NC)

    def set_state(self, q: np.ndarray) -> None:
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        L.check(L.lib().mxm_dg_upload_q(self._h, q.ctypes.data, q.size), "mxm_dg_upload_q")

    def get_state(self) -> np.ndarray:
        q = np.zeros(self.shape)
        L.check(L.lib().mxm_dg_download_q(self._h, q.ctypes.data, q.size), "mxm_dg_download_q")
        return q

    def download_owned_component(self, c: int) -> np.ndarray:
        """Component c at every node, [K*NP] (src/solver.mojo:711-729)."""
        out = np.zeros(self.num_elements * self.NP)
        L.check(L.lib().mxm_dg_download_component(self._h, c, out.ctypes.data, out.size),
                "mxm_dg_download_component")
        return out

    # ---- stepping ------------------------------------------------------------
    def rk_stage(self, stage: int, dt: float) -> None:
        L.check(L.lib().mxm_dg_rk_stage(self._h, stage, dt), "mxm_dg_rk_stage")

    def step_ssprk3(self, dt: float, nsteps: int = 1) -> None:
        L.check(L.lib().mxm_dg_step_ssprk3(self._h, dt, nsteps), "mxm_dg_step_ssprk3")

    def prepare_steps(self, dt: float, nsteps: int) -> None:
        """Build the replayed graphs step_ssprk3(dt, nsteps) uses, without
        stepping (keeps graph construction out of a timed region)."""
        L.check(L.lib().mxm_dg_prepare_steps(self._h, dt, nsteps), "mxm_dg_prepare_steps")

    def enable_cell_limiter(self, enabled: bool = True, venkat_eps: float = 0.1) -> None:
        L.check(L.lib().mxm_dg_enable_limiter(self._h, int(enabled), venkat_eps), "mxm_dg_enable_limiter")

    def synchronize(self) -> None:
        L.check(L.lib().mxm_dg_sync(self._h), "mxm_dg_sync")

    # ---- reports --------------------------------------------------------------
    def dof_count(self) -> int:
        v = C.c_longlong()
        L.check(L.lib().mxm_dg_dof_count(self._h, C.byref(v)), "mxm_dg_dof_count")
        return v.value

    def state_bytes_per_step(self) -> int:
        v = C.c_longlong()
        L.check(L.lib().mxm_dg_state_bytes_per_step(self._h, C.byref(v)), "mxm_dg_state_bytes_per_step")
        return v.value

    def bench_step_loop(self, dt: float, warmup_steps: int = 5, measure_steps: int = 50) -> ThroughputReport:
        t = L.Throughput()
        L.check(L.lib().mxm_dg_bench_step_loop(self._h, dt, warmup_steps, measure_steps, C.byref(t)),
                "mxm_dg_bench_step_loop")
        return ThroughputReport(t.steps, t.wall_seconds, t.dof, t.state_bytes_per_step, t.stage_seconds,
                                t.achieved_hbm_gbs, t.hbm_fraction, t.mfma_tflops, t.mfma_utilization)

    def memory_report(self) -> MemoryReport:
        m = L.MemReport()
        L.check(L.lib().mxm_dg_memory_report(self._h, C.byref(m)), "mxm_dg_memory_report")
        return MemoryReport(m.rk_stage, m.dg_operators, m.limiter, m.mesh_connectivity, m.halo_device, m.halo_pinned)

    def diagnostics(self):
        """Mass-weighted int q_c dV and int q_c^2 dV, max |q_c|, kinetic
        energy and enstrophy of the current state, computed on the device
        in compensated FP64 (mxm_dg_diagnostics)."""
        from .diagnostics import Diagnostics
        nc = self.NC
        out = (C.c_double * (3 * nc + 2))()
        n = C.c_int()
        L.check(L.lib().mxm_dg_diagnostics(self._h, out, 3 * nc + 2, C.byref(n)), "mxm_dg_diagnostics")
        return Diagnostics.from_flat(list(out), nc)

    def max_wave_speed(self) -> float:
        """Largest signal speed over the nodes of the current state, reduced
        on the device (mxm_dg_max_wave_speed); dt = CFL h / (speed (2P+1))
        is time_integrator.cfl_dt."""
        v = C.c_double()
        L.check(L.lib().mxm_dg_max_wave_speed(self._h, C.byref(v)), "mxm_dg_max_wave_speed")
        return v.value

    def max_wave_speed_noalloc(self, speed_dev_ptr: int) -> None:
        """Enqueue the same reduction with no allocation
        (mxm_dg_max_wave_speed_noalloc): the speed goes to the one double of
        device memory at speed_dev_ptr once the solver's stream reaches it
        (synchronize()); the tile partials use the library's fixed scratch
        array (meshes of at most 262144 tiles)."""
        L.check(L.lib().mxm_dg_max_wave_speed_noalloc(self._h, C.c_void_p(speed_dev_ptr)),
                "mxm_dg_max_wave_speed_noalloc")

    def autotune(self, dt: float, reps: int = 20):
        """Time every launch configuration of the stage kernel on this mesh
        (reps SSPRK3 steps of the current state per configuration, after
        untimed warm-up steps; the state is saved and restored) and keep the fastest; for a shape with a
        wide tile (`tile`), both tile sizes are timed and the faster tiling
        kept.  Returns (chosen index, microseconds per stage of each
        configuration of the kept tiling)."""
        us = (C.c_double * 4)()
        cfg = C.c_int()
        L.check(L.lib().mxm_dg_autotune(self._h, dt, reps, C.byref(cfg), us, 4), "mxm_dg_autotune")
        return cfg.value, [us[k] for k in range(launch_config_count(self.precision))]

    @property
    def launch_config(self) -> int:
        v = C.c_int()
        L.check(L.lib().mxm_dg_launch_config(self._h, -1, C.byref(v)), "mxm_dg_launch_config")
        return v.value

    @launch_config.setter
    def launch_config(self, cfg: int) -> None:
        L.check(L.lib().mxm_dg_launch_config(self._h, int(cfg), None), "mxm_dg_launch_config")

    @property
    def tile(self) -> tuple:
        """(wide, elements per tile, threads per block) of the stage kernel's
        tiling; wide tiles exist for the FP32-MFMA form on triangles at
        P = 3."""
        e, t = C.c_int(), C.c_int()
        L.check(L.lib().mxm_dg_tile_config(self._h, -1, C.byref(e), C.byref(t)), "mxm_dg_tile_config")
        return (t.value > 256, e.value, t.value)

    @tile.setter
    def tile(self, wide: bool) -> None:
        """Lay the state out in wide tiles (True) or 256-thread tiles; the
        state keeps its values."""
        L.check(L.lib().mxm_dg_tile_config(self._h, 1 if wide else 0, None, None), "mxm_dg_tile_config")

    def time_stage_kernel(self, dt: float, nsteps: int) -> float:
        """Average stage-kernel duration in microseconds (HIP events on the
        solver's stream around nsteps steps, limiter off for them and
        restored after), after about 30 ms of untimed warm-up steps and the
        same nsteps steps once; advances the state by all of them."""
        v = C.c_double()
        L.check(L.lib().mxm_dg_time_stage_kernel(self._h, dt, nsteps, C.byref(v)), "mxm_dg_time_stage_kernel")
        return v.value
