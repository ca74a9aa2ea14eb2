"""The C ABI of libmxm_hip.so without a GPU: the library loads, exports every
function include/mxm_dg.h declares (and the Python binding binds exactly
that set), the host-only entry points follow the reference shim's
conventions -- int return codes, MXM_OK = 0, a message from
mxm_last_error() (src/mpi_shim.c:41-177, src/mpi.mojo:35-37) -- and no
kernel is launched."""
import ctypes as C
import os
import re

import numpy as np

from mojoxm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mxm_dg.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*(?:int|const char\s*\*)\s*(mxm_\w+)\s*\(", text, flags=re.M))


def declared_constants():
    text = open(HEADER).read()
    return {k: int(v) for k, v in re.findall(r"#define\s+(MXM_\w+)\s+(-?\d+)\b", text)}


def test_header_parses():
    names = declared_functions()
    assert "mxm_hip_init" in names and "mxm_dg_step_ssprk3" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    L = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    assert set(_lib.EXPORTS) == declared_functions()


def test_constants_match_header():
    k = declared_constants()
    assert k["MXM_OK"] == _lib.MXM_OK == 0
    assert k["MXM_PRM_MAX"] == _lib.PRM_MAX
    # BC values of src/boundary.mojo:24-38
    assert (k["MXM_BC_INTERIOR"], k["MXM_BC_WALL"], k["MXM_BC_OUTFLOW"], k["MXM_BC_INFLOW"]) == (0, 1, 2, 3)
    v = C.c_int()
    assert _lib.lib().mxm_get_constant(0, C.byref(v)) == 0 and v.value == k["MXM_PRM_MAX"]
    assert _lib.lib().mxm_get_constant(1, C.byref(v)) == 0 and v.value > 0
    assert _lib.lib().mxm_get_constant(99, C.byref(v)) != 0
    assert _lib.lib().mxm_last_error()


def test_errors_are_codes_with_messages():
    L = _lib.lib()
    # null handles and bad arguments are rejected before any device work
    assert L.mxm_dg_info(None, (C.c_int * 5)()) != 0
    assert b"" != L.mxm_last_error()
    assert L.mxm_mesh_info(None, (C.c_int * 6)()) != 0
    D = np.zeros(3 * 20 * 20)
    assert L.mxm_refelem_3d(0, D.ctypes.data, None, None, None) != 0
    assert L.mxm_refelem_3d(99, D.ctypes.data, None, None, None) != 0
    msg = L.mxm_last_error().decode()
    assert msg
    d = _lib.Mesh2DDesc(nx=0, ny=4, lx=1.0, ly=1.0)
    h = C.c_void_p()
    assert L.mxm_mesh_create_2d(C.byref(d), 2, C.byref(h)) != 0
    bad_bc = _lib.Mesh2DDesc(nx=4, ny=4, lx=1.0, ly=1.0)
    bad_bc.bc[0] = 7
    assert L.mxm_mesh_create_2d(C.byref(bad_bc), 2, C.byref(h)) != 0


def test_host_mesh_entry_points_round_trip():
    """mxm_mesh_create_2d / info / get / destroy run on the host only."""
    L = _lib.lib()
    d = _lib.Mesh2DDesc(nx=3, ny=2, lx=1.5, ly=1.0)
    h = C.c_void_p()
    assert L.mxm_mesh_create_2d(C.byref(d), 2, C.byref(h)) == 0
    info = (C.c_int * 6)()
    assert L.mxm_mesh_info(h, info) == 0
    dim, K, Nf, NP, NFP, NF = info
    assert (dim, K, Nf, NP, NFP, NF) == (2, 12, 18, 6, 3, 3)
    faces = np.zeros(K * 3, np.int32)
    assert L.mxm_mesh_get(h, _lib.T_ELEM_FACES, faces.ctypes.data) == 0
    assert faces.min() >= 0 and faces.max() < Nf
    assert L.mxm_mesh_get(h, 99, faces.ctypes.data) != 0
    assert L.mxm_mesh_destroy(h) == 0
