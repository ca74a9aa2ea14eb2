"""Reference-element operators (ReferenceElement[P], src/reference.mojo:519-660;
ReferenceElement2D[P], src/reference_2d.mojo:130-261) built by the library's
host code."""
from __future__ import annotations

import numpy as np

from ._lib import check, lib


def num_tet_nodes(P: int) -> int:
    return (P + 1) * (P + 2) * (P + 3) // 6


def num_tri_nodes(P: int) -> int:
    return (P + 1) * (P + 2) // 2


class ReferenceElement:
    """Tetrahedron of order P: D_ref[3,NP,NP], Lift_ref[4,NP,NFP],
    node_weights[NP], face_to_elem[4,NFP]."""

    def __init__(self, P: int = 2):
        self.P = P
        NP, NFP = num_tet_nodes(P), num_tri_nodes(P)
        self.D_ref = np.zeros((3, NP, NP))
        self.Lift_ref = np.zeros((4, NP, NFP))
        self.node_weights = np.zeros(NP)
        self.face_to_elem = np.zeros((4, NFP), np.int32)
        check(lib().mxm_refelem_3d(P, self.D_ref.ctypes.data, self.Lift_ref.ctypes.data,
                                   self.node_weights.ctypes.data, self.face_to_elem.ctypes.data), "mxm_refelem_3d")
        # nodal differentiation d phi_j / d r_k at node i (diagnostics)
        self.D_nodal = np.zeros((3, NP, NP))
        check(lib().mxm_refelem_deriv(3, P, self.D_nodal.ctypes.data), "mxm_refelem_deriv")


class ReferenceElement2D:
    """Triangle of order P: D_ref[2,NP,NP], Lift_ref[3,NP,P+1],
    node_weights[NP], edge_to_elem[3,P+1]."""

    def __init__(self, P: int = 2):
        self.P = P
        NP, NE = num_tri_nodes(P), P + 1
        self.D_ref = np.zeros((2, NP, NP))
        self.Lift_ref = np.zeros((3, NP, NE))
        self.node_weights = np.zeros(NP)
        self.edge_to_elem = np.zeros((3, NE), np.int32)
        check(lib().mxm_refelem_2d(P, self.D_ref.ctypes.data, self.Lift_ref.ctypes.data,
                                   self.node_weights.ctypes.data, self.edge_to_elem.ctypes.data), "mxm_refelem_2d")
        self.D_nodal = np.zeros((2, NP, NP))
        check(lib().mxm_refelem_deriv(2, P, self.D_nodal.ctypes.data), "mxm_refelem_deriv")
