"""Host-built meshes in the reference's table layout: LocalMesh[P]
(src/local_mesh.mojo:1195-1488) and LocalMesh2D[P]
(src/local_mesh_2d.mojo:73-395)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .boundary import BoundaryConditions, BoundaryConditions2D


class _MeshBase:
    _TABLES = {
        "elem_node_xyz": (L.T_ELEM_NODE_XYZ, np.float64), "elem_invJ": (L.T_ELEM_INVJ, np.float64),
        "elem_inv_vol": (L.T_ELEM_INV_VOL, np.float64), "elem_faces": (L.T_ELEM_FACES, np.int32),
        "elem_face_side": (L.T_ELEM_FACE_SIDE, np.int32), "elem_canon_to_ref": (L.T_ELEM_CANON_TO_REF, np.int32),
        "face_elem": (L.T_FACE_ELEM, np.int32), "face_elem_node": (L.T_FACE_ELEM_NODE, np.int32),
        "face_normal": (L.T_FACE_NORMAL, np.float64), "face_area": (L.T_FACE_AREA, np.float64),
        "face_bc_type": (L.T_FACE_BC_TYPE, np.int32),
    }

    def _load(self, handle):
        info = (C.c_int * 6)()
        L.check(L.lib().mxm_mesh_info(handle, info), "mxm_mesh_info")
        self.ndim, self.num_elements, self.num_faces, self.NP, self.NFP, self.NF = list(info)
        self.K, self.Nf = self.num_elements, self.num_faces
        K, Nf, NP, NFP, NF, d = self.num_elements, self.num_faces, self.NP, self.NFP, self.NF, self.ndim
        shapes = {
            "elem_node_xyz": (K, NP, d), "elem_invJ": (K, d, d), "elem_inv_vol": (K,), "elem_faces": (K, NF),
            "elem_face_side": (K, NF), "elem_canon_to_ref": (K, NF, NFP), "face_elem": (Nf, 2),
            "face_elem_node": (Nf, 2, NFP), "face_normal": (Nf, d), "face_area": (Nf,), "face_bc_type": (Nf,),
        }
        try:
            for name, (tid, dt) in self._TABLES.items():
                a = np.zeros(shapes[name], dtype=dt)
                L.check(L.lib().mxm_mesh_get(handle, tid, a.ctypes.data), "mxm_mesh_get")
                setattr(self, name, a)
        finally:
            L.lib().mxm_mesh_destroy(handle)


class LocalMesh(_MeshBase):
    """Periodic Kuhn-tet mesh of (nx, ny, nz) cubes, 6 tets each."""

    def __init__(self, nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, bcs: BoundaryConditions | None = None, P: int = 2):
        bcs = bcs or BoundaryConditions.periodic()
        d = L.Mesh3DDesc(nx, ny, nz, lx, ly, lz, (C.c_int * 6)(*bcs.as_list()))
        h = C.c_void_p()
        L.check(L.lib().mxm_mesh_create_3d(C.byref(d), P, C.byref(h)), "mxm_mesh_create_3d")
        self.P, self.n, self.L, self.bcs = P, (nx, ny, nz), (lx, ly, lz), bcs
        self._load(h)


class LocalMesh2D(_MeshBase):
    """/-split triangle mesh of (nx, ny) squares, 2 triangles each."""

    def __init__(self, nx, ny, lx=1.0, ly=1.0, bcs: BoundaryConditions2D | None = None, P: int = 2):
        bcs = bcs or BoundaryConditions2D.periodic()
        d = L.Mesh2DDesc(nx, ny, lx, ly, (C.c_int * 4)(*bcs.as_list()))
        h = C.c_void_p()
        L.check(L.lib().mxm_mesh_create_2d(C.byref(d), P, C.byref(h)), "mxm_mesh_create_2d")
        self.P, self.n, self.L, self.bcs = P, (nx, ny), (lx, ly), bcs
        self._load(h)
