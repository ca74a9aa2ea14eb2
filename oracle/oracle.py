"""ctypes front end of the CPU FP64 oracle (test infrastructure).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module; the shipped library never does.  It loads
``oracle/liboracle.so`` (built by ``oracle/Makefile``) and exposes the
restated reference routines on numpy arrays.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

(PHYS_ADVECTION, PHYS_EULER, PHYS_EULER2D, PHYS_SW2D, PHYS_MHDGLM2D, PHYS_MAXWELL2D, PHYS_ADVECTION2D, PHYS_SW, PHYS_MAXWELL,
 PHYS_MHD, PHYS_TWOFLUID) = range(11)

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class _Mesh(C.Structure):
    _fields_ = [
        ("ndim", C.c_int), ("P", C.c_int), ("NP", C.c_int), ("NFP", C.c_int), ("NF", C.c_int),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("num_elements", C.c_int), ("num_faces", C.c_int),
        ("elem_node_xyz", C.c_void_p), ("elem_invJ", C.c_void_p), ("elem_inv_vol", C.c_void_p),
        ("elem_faces", C.c_void_p), ("elem_face_side", C.c_void_p), ("elem_canon_to_ref", C.c_void_p),
        ("face_elem", C.c_void_p), ("face_elem_node", C.c_void_p), ("face_normal", C.c_void_p),
        ("face_area", C.c_void_p), ("face_bc_type", C.c_void_p),
    ]


class _Ops(C.Structure):
    _fields_ = [
        ("ndim", C.c_int), ("P", C.c_int), ("NP", C.c_int), ("NFP", C.c_int), ("NF", C.c_int),
        ("D", C.c_void_p), ("Lift", C.c_void_p), ("w", C.c_void_p),
    ]


def build(force: bool = False) -> None:
    """Compile liboracle.so from the sources beside this file."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_refelem3d.argtypes = [C.c_int, _dp, _dp, _dp, _ip, _dp]
        L.orc_refelem2d.argtypes = [C.c_int, _dp, _dp, _dp, _ip, _dp]
        L.orc_mat_inv.argtypes = [_dp, C.c_int, _dp]
        L.orc_mesh3d_counts.argtypes = [C.c_int, C.c_int, C.c_int, _ip, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_mesh2d_counts.argtypes = [C.c_int, C.c_int, _ip, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_mesh3d_build.argtypes = [C.POINTER(_Mesh), _ip]
        L.orc_mesh2d_build.argtypes = [C.POINTER(_Mesh), _ip]
        L.orc_phys_nc.argtypes = [C.c_int]
        L.orc_internal_flux.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_numerical_flux.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double, _dp, C.POINTER(C.c_double)]
        L.orc_boundary_flux.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_double, C.c_double, _dp, C.POINTER(C.c_double)]
        L.orc_limit_state.argtypes = [C.c_int, _dp, _dp]
        L.orc_face_flux_2d.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, C.c_double, C.c_double, _dp]
        L.orc_stage.argtypes = [C.POINTER(_Mesh), C.POINTER(_Ops), C.c_int, _dp, _dp, _dp, _dp, _dp,
                                C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_limiter.argtypes = [C.POINTER(_Mesh), C.POINTER(_Ops), C.c_int, _dp, C.c_double]
        L.orc_step_ssprk3.argtypes = [C.POINTER(_Mesh), C.POINTER(_Ops), C.c_int, _dp, _dp, C.c_double,
                                      C.c_int, C.c_int, C.c_double]
        L.orc_nodal_deriv.argtypes = [C.c_int, C.c_int, _dp]
        L.orc3_source.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_diagnostics.argtypes = [C.POINTER(_Mesh), _dp, _dp, C.c_int, _dp, _dp]
        L.orc_max_speed.argtypes = [C.c_int, _dp, _dp]
        L.orc_max_speed.restype = C.c_double
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        _lib = L
    return _lib


def nt(P: int) -> int:
    return (P + 1) * (P + 2) * (P + 3) // 6


def ntri(P: int) -> int:
    return (P + 1) * (P + 2) // 2


def refelem3d(P: int) -> dict:
    NP, NFP = nt(P), ntri(P)
    out = dict(D=np.zeros(3 * NP * NP), Lift=np.zeros(4 * NP * NFP), w=np.zeros(NP),
               f2e=np.zeros(4 * NFP, np.int32), node_pos=np.zeros(NP * 3))
    rc = lib().orc_refelem3d(P, out["D"], out["Lift"], out["w"], out["f2e"], out["node_pos"])
    if rc:
        raise RuntimeError(f"orc_refelem3d rc={rc}")
    out["D"] = out["D"].reshape(3, NP, NP)
    out["Lift"] = out["Lift"].reshape(4, NP, NFP)
    out["f2e"] = out["f2e"].reshape(4, NFP)
    out["node_pos"] = out["node_pos"].reshape(NP, 3)
    return out


def refelem2d(P: int) -> dict:
    NP, NE = ntri(P), P + 1
    out = dict(D=np.zeros(2 * NP * NP), Lift=np.zeros(3 * NP * NE), w=np.zeros(NP),
               e2e=np.zeros(3 * NE, np.int32), node_pos=np.zeros(NP * 2))
    rc = lib().orc_refelem2d(P, out["D"], out["Lift"], out["w"], out["e2e"], out["node_pos"])
    if rc:
        raise RuntimeError(f"orc_refelem2d rc={rc}")
    out["D"] = out["D"].reshape(2, NP, NP)
    out["Lift"] = out["Lift"].reshape(3, NP, NE)
    out["e2e"] = out["e2e"].reshape(3, NE)
    out["node_pos"] = out["node_pos"].reshape(NP, 2)
    return out


def mat_inv(m: np.ndarray) -> np.ndarray:
    m = np.ascontiguousarray(m, dtype=np.float64)
    out = np.zeros_like(m)
    if lib().orc_mat_inv(m.ravel(), m.shape[0], out.ravel()):
        raise ZeroDivisionError("singular")
    return out


class Mesh:
    """A mesh in the reference's table layout, built by the oracle."""

    NAMES = ("elem_node_xyz", "elem_invJ", "elem_inv_vol", "elem_faces", "elem_face_side",
             "elem_canon_to_ref", "face_elem", "face_elem_node", "face_normal", "face_area",
             "face_bc_type")

    def __init__(self, ndim, P, n, L, bc=None):
        self.ndim, self.P = ndim, P
        if ndim == 3:
            bc = np.zeros(6, np.int32) if bc is None else np.asarray(bc, np.int32)
            ne, nf = C.c_int(), C.c_int()
            lib().orc_mesh3d_counts(n[0], n[1], n[2], bc, C.byref(ne), C.byref(nf))
            NP, NFP, NF = nt(P), ntri(P), 4
        else:
            bc = np.zeros(4, np.int32) if bc is None else np.asarray(bc, np.int32)
            ne, nf = C.c_int(), C.c_int()
            lib().orc_mesh2d_counts(n[0], n[1], bc, C.byref(ne), C.byref(nf))
            NP, NFP, NF = ntri(P), P + 1, 3
        self.bc = bc
        K, Nf = ne.value, nf.value
        self.K, self.Nf, self.NP, self.NFP, self.NF = K, Nf, NP, NFP, NF
        self.num_elements, self.num_faces = K, Nf
        d = ndim
        self.elem_node_xyz = np.zeros((K, NP, d))
        self.elem_invJ = np.zeros((K, d, d))
        self.elem_inv_vol = np.zeros(K)
        self.elem_faces = np.zeros((K, NF), np.int32)
        self.elem_face_side = np.zeros((K, NF), np.int32)
        self.elem_canon_to_ref = np.zeros((K, NF, NFP), np.int32)
        self.face_elem = np.zeros((Nf, 2), np.int32)
        self.face_elem_node = np.zeros((Nf, 2, NFP), np.int32)
        self.face_normal = np.zeros((Nf, d))
        self.face_area = np.zeros(Nf)
        self.face_bc_type = np.zeros(Nf, np.int32)
        s = _Mesh()
        s.ndim, s.P = ndim, P
        s.nx, s.ny = n[0], n[1]
        s.nz = n[2] if ndim == 3 else 1
        s.lx, s.ly = L[0], L[1]
        s.lz = L[2] if ndim == 3 else 1.0
        for name in self.NAMES:
            setattr(s, name, getattr(self, name).ctypes.data)
        self._s = s
        rc = (lib().orc_mesh3d_build if ndim == 3 else lib().orc_mesh2d_build)(C.byref(s), bc)
        if rc:
            raise RuntimeError(f"oracle mesh build rc={rc}")
        self.n, self.L = tuple(n), tuple(L)
        re = refelem3d(P) if ndim == 3 else refelem2d(P)
        self.ref = re
        o = _Ops()
        o.ndim, o.P, o.NP, o.NFP, o.NF = ndim, P, NP, NFP, NF
        self._D = np.ascontiguousarray(re["D"].ravel())
        self._Lift = np.ascontiguousarray(re["Lift"].ravel())
        self._w = np.ascontiguousarray(re["w"])
        o.D, o.Lift, o.w = self._D.ctypes.data, self._Lift.ctypes.data, self._w.ctypes.data
        self._o = o

    def stage(self, phys, prm, q_in, q_a, q_b, a, b, c, dt):
        q_out = np.zeros_like(q_in)
        rc = lib().orc_stage(C.byref(self._s), C.byref(self._o), phys, _prm(prm), _f(q_in), _f(q_a),
                             _f(q_b), q_out.reshape(-1), a, b, c, dt)
        if rc:
            raise RuntimeError(f"orc_stage rc={rc}")
        return q_out

    def step(self, phys, prm, q, dt, nsteps, limiter=False, eps=0.1):
        q = np.array(q, dtype=np.float64, copy=True)
        rc = lib().orc_step_ssprk3(C.byref(self._s), C.byref(self._o), phys, _prm(prm), q.reshape(-1),
                                   dt, nsteps, int(limiter), eps)
        if rc:
            raise RuntimeError(f"orc_step_ssprk3 rc={rc}")
        return q

    def diagnostics(self, phys, q):
        """Mass-weighted int q_c dV, int q_c^2 dV, max |q_c|, KE, enstrophy
        (diag.c); returns the flat array of 3*NC + 2 values."""
        nc = lib().orc_phys_nc(phys)
        if not hasattr(self, "_Dn"):
            self._Dn = nodal_deriv(self.ndim, self.P).ravel()
        out = np.zeros(3 * nc + 2)
        rc = lib().orc_diagnostics(C.byref(self._s), self._w, self._Dn, phys, _f(q), out)
        if rc:
            raise RuntimeError(f"orc_diagnostics rc={rc}")
        return out

    def limiter(self, nc, q, eps=0.1):
        q = np.array(q, dtype=np.float64, copy=True)
        lib().orc_limiter(C.byref(self._s), C.byref(self._o), nc, q.reshape(-1), eps)
        return q


def nodal_deriv(ndim: int, P: int) -> np.ndarray:
    """Dn[k][i][j] = d phi_j / d r_k at node i."""
    NP = nt(P) if ndim == 3 else ntri(P)
    Dn = np.zeros((ndim, NP, NP))
    if lib().orc_nodal_deriv(ndim, P, Dn.reshape(-1)):
        raise RuntimeError("orc_nodal_deriv failed")
    return Dn


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(-1)


def _prm(p):
    out = np.zeros(32)
    p = np.asarray(p, dtype=np.float64).ravel()
    out[: p.size] = p
    return out


def numerical_flux(phys, prm, ql, qr, n):
    nc = lib().orc_phys_nc(phys)
    f = np.zeros(nc)
    sp = C.c_double()
    n3 = list(n) + [0.0] * (3 - len(n))
    rc = lib().orc_numerical_flux(phys, _prm(prm), _f(ql), _f(qr), n3[0], n3[1], n3[2], f, C.byref(sp))
    if rc:
        raise RuntimeError("numerical_flux")
    return f, sp.value


def boundary_flux(phys, prm, q, bc, n):
    nc = lib().orc_phys_nc(phys)
    f = np.zeros(nc)
    sp = C.c_double()
    n3 = list(n) + [0.0] * (3 - len(n))
    rc = lib().orc_boundary_flux(phys, _prm(prm), _f(q), bc, n3[0], n3[1], n3[2], f, C.byref(sp))
    if rc:
        raise RuntimeError("boundary_flux")
    return f, sp.value


def internal_flux(phys, prm, q):
    nc = lib().orc_phys_nc(phys)
    f = np.zeros(3 * nc)
    if lib().orc_internal_flux(phys, _prm(prm), _f(q), f):
        raise RuntimeError("internal_flux")
    return f.reshape(3, nc)


def max_speed(phys, prm, q):
    """Largest signal speed of one node state (cfl.c)."""
    return lib().orc_max_speed(phys, _prm(prm), _f(q))


def max_wave_speed(phys, prm, q):
    """Largest signal speed over every node of a state [K, NP, NC]; NaN if
    any node's is."""
    q = np.asarray(q, dtype=np.float64)
    nc = q.shape[-1]
    sp = [max_speed(phys, prm, v) for v in q.reshape(-1, nc)]
    return float(np.nan if np.isnan(sp).any() else max(sp))


def source(phys, prm, q):
    """source_term of the other trait physics (physics3d.c)."""
    nc = lib().orc_phys_nc(phys)
    s = np.zeros(nc)
    if lib().orc3_source(phys, _prm(prm), _f(q), s):
        raise RuntimeError("source")
    return s


def limit_state(phys, prm, q):
    q = np.array(q, dtype=np.float64, copy=True)
    lib().orc_limit_state(phys, _prm(prm), q)
    return q


def face_flux_2d(phys, prm, ql, qr, bc, n):
    nc = lib().orc_phys_nc(phys)
    f = np.zeros(nc)
    if lib().orc_face_flux_2d(phys, _prm(prm), _f(ql), _f(qr), bc, n[0], n[1], f):
        raise RuntimeError("face_flux_2d")
    return f


def set_threads(n: int) -> None:
    lib().orc_set_threads(n)


def get_threads() -> int:
    return lib().orc_get_threads()
