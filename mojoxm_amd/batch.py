"""Batched solves over the C ABI of libmxm_hip.so (mxm_batch_*): many
independent problems advanced together, one stage launch per kernel group
and RK stage, one replayed HIP graph per SSPRK3 step.

No counterpart in the reference, whose drivers advance one problem per
Solver (e.g. examples/mhd_alfven_glm_2d_gpu.mojo:215-242); every instance
follows Solver.step_ssprk3 exactly and its result does not depend on the
other instances in the batch."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .boundary import BoundaryConditions, BoundaryConditions2D
from .solver import PREC_FP64


def _desc(n, L_box, bcs):
    if len(n) == 3:
        bcs = bcs or BoundaryConditions.periodic()
        return L.Mesh3DDesc(n[0], n[1], n[2], L_box[0], L_box[1], L_box[2], (C.c_int * 6)(*bcs.as_list()))
    bcs = bcs or BoundaryConditions2D.periodic()
    return L.Mesh2DDesc(n[0], n[1], L_box[0], L_box[1], (C.c_int * 4)(*bcs.as_list()))


def _nodes(ndim, P):
    return (P + 1) * (P + 2) * (P + 3) // 6 if ndim == 3 else (P + 1) * (P + 2) // 2


def footprint(n, L_box, physics, P, precision=PREC_FP64, limiter=False, bcs=None) -> int:
    """Exact device bytes one instance adds to a batch (host only)."""
    d = _desc(n, L_box, bcs)
    out = C.c_longlong()
    fn = L.lib().mxm_batch_footprint_3d if len(n) == 3 else L.lib().mxm_batch_footprint_2d
    L.check(fn(C.byref(d), P, physics.phys_id, precision, int(limiter), C.byref(out)), "mxm_batch_footprint")
    return out.value


class Batch:
    """A set of independent instances advanced together on one GPU."""

    def __init__(self, precision: int = PREC_FP64, device: int = 0):
        L.ensure_device(device)
        self.precision = precision
        h = C.c_void_p()
        L.check(L.lib().mxm_batch_create(precision, C.byref(h)), "mxm_batch_create")
        self._h = h
        self._shape = {}

    def close(self) -> None:
        if getattr(self, "_h", None):
            L.lib().mxm_batch_destroy(self._h)
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

    def autotune(self, on: bool = True) -> None:
        """Autotune the stage kernels' launch configuration per kernel group
        and tile-count range when groups are packed (on by default)."""
        L.check(L.lib().mxm_batch_autotune(self._h, int(bool(on))), "mxm_batch_autotune")

    def graphs(self, on: bool = True) -> None:
        """Replay each chunk of steps as a HIP graph (default), or launch
        the same kernels directly on the streams (for kernel tracers)."""
        L.check(L.lib().mxm_batch_graphs(self._h, int(bool(on))), "mxm_batch_graphs")

    def headroom(self, fraction: float) -> None:
        """Spare capacity (a fraction of tiles and slots) for arenas that are
        rebuilt because a new instance did not fit; 0 (default) is exact."""
        L.check(L.lib().mxm_batch_headroom(self._h, float(fraction)), "mxm_batch_headroom")

    def add(self, n, L_box, physics, P, q0, dt, nsteps, bcs=None, limiter=False, eps=0.1) -> int:
        """Queue one instance; it joins at the next step().  Returns its id."""
        ndim = len(n)
        K = int(np.prod(n)) * (6 if ndim == 3 else 2)
        shape = (K, _nodes(ndim, P), physics.NUM_COMPONENTS)
        q0 = np.ascontiguousarray(q0, dtype=np.float64).reshape(-1)
        if q0.size != np.prod(shape):
            raise ValueError(f"initial state must have shape {shape}")
        prm = np.ascontiguousarray(physics.params(), dtype=np.float64)
        d = _desc(n, L_box, bcs)
        ident = C.c_int()
        fn = L.lib().mxm_batch_add_3d if ndim == 3 else L.lib().mxm_batch_add_2d
        L.check(fn(self._h, C.byref(d), P, physics.phys_id, prm.ctypes.data, float(dt), int(nsteps), int(limiter),
                   float(eps), q0.ctypes.data, q0.size, C.byref(ident)), "mxm_batch_add")
        self._shape[ident.value] = shape
        return ident.value

    def step(self, nsteps: int = 1) -> None:
        L.check(L.lib().mxm_batch_step(self._h, int(nsteps)), "mxm_batch_step")

    def poll(self, cap: int = 4096, lagged: bool = False) -> list:
        """Ids of instances that finished (or failed) since the last poll.
        Synchronises; with lagged=True it waits only for the work of the
        step() calls before the most recent one, which keeps running."""
        ids = (C.c_int * cap)()
        cnt = C.c_int()
        fn = L.lib().mxm_batch_poll_lagged if lagged else L.lib().mxm_batch_poll
        L.check(fn(self._h, ids, cap, C.byref(cnt)), "mxm_batch_poll")
        return list(ids[: cnt.value])

    def status(self, ident: int) -> int:
        s = C.c_int()
        L.check(L.lib().mxm_batch_status(self._h, ident, C.byref(s)), "mxm_batch_status")
        return s.value

    def failed(self, ident: int) -> bool:
        """True when a stored value of the instance became non-finite (it
        was switched off at that step; poll reports it)."""
        f = C.c_int()
        L.check(L.lib().mxm_batch_failed(self._h, ident, C.byref(f)), "mxm_batch_failed")
        return bool(f.value)

    def get_state(self, ident: int) -> np.ndarray:
        q = np.zeros(self._shape[ident])
        L.check(L.lib().mxm_batch_get_state(self._h, ident, q.ctypes.data, q.size), "mxm_batch_get_state")
        return q

    def diagnostics(self, ident: int):
        """mxm_dg_diagnostics of one instance."""
        from .diagnostics import Diagnostics
        nc = self._shape[ident][2]
        out = (C.c_double * (3 * nc + 2))()
        n = C.c_int()
        L.check(L.lib().mxm_batch_diagnostics(self._h, ident, out, 3 * nc + 2, C.byref(n)), "mxm_batch_diagnostics")
        return Diagnostics.from_flat(list(out), nc)

    def max_wave_speed(self, ident: int) -> float:
        """mxm_dg_max_wave_speed of one instance."""
        v = C.c_double()
        L.check(L.lib().mxm_batch_max_wave_speed(self._h, ident, C.byref(v)), "mxm_batch_max_wave_speed")
        return v.value

    def remove(self, ident: int) -> None:
        L.check(L.lib().mxm_batch_remove(self._h, ident), "mxm_batch_remove")
        self._shape.pop(ident, None)

    def memory(self) -> int:
        b = C.c_longlong()
        L.check(L.lib().mxm_batch_memory(self._h, C.byref(b)), "mxm_batch_memory")
        return b.value

    def synchronize(self) -> None:
        L.check(L.lib().mxm_batch_sync(self._h), "mxm_batch_sync")
