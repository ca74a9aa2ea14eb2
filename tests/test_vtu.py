"""VTU / PVD output (mojoxm_amd.vtu): a round trip of points, cells and
fields through the appended-binary format, the cell types, and the element
node order against VTK's quadratic cells and Lagrange edge order."""
import os

import numpy as np
import pytest

import mojoxm_amd as M
from mojoxm_amd import vtu

EDGES3 = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
EDGES2 = [(0, 1), (1, 2), (2, 0)]


@pytest.mark.parametrize("dim,P,ctype", [(2, 1, 5), (2, 2, 22), (2, 3, 69), (3, 1, 71), (3, 2, 24), (3, 3, 71)])
def test_round_trip(tmp_path, dim, P, ctype):
    m = M.LocalMesh(2, 2, 2, 1.0, 1.0, 0.5, P=P) if dim == 3 else M.LocalMesh2D(3, 2, 1.0, 0.5, P=P)
    X = m.elem_node_xyz
    K, NP = X.shape[:2]
    rng = np.random.default_rng(P)
    f = {"rho": rng.standard_normal((K, NP)), "p": rng.standard_normal((K, NP))}
    path = str(tmp_path / "a.vtu")
    vtu.write_vtu(path, X, f)
    r = vtu.read_vtu(path)
    pts = np.zeros((K * NP, 3))
    pts[:, :dim] = X.reshape(-1, dim)
    assert np.array_equal(r["points"], pts.astype(np.float32))
    assert np.array_equal(r["connectivity"], np.arange(K * NP))
    assert np.array_equal(r["offsets"], NP * np.arange(1, K + 1))
    assert np.all(r["types"] == ctype)
    for k in f:
        assert np.array_equal(r["fields"][k], f[k].astype(np.float32).ravel())
    assert open(path, "rb").read().startswith(b'<?xml version="1.0"?>')


def test_quadratic_node_order_is_vtk():
    X = M.LocalMesh(2, 2, 2, 1.0, 1.0, 1.0, P=2).elem_node_xyz
    for e in range(X.shape[0]):
        for k, (a, b) in enumerate(EDGES3):
            assert np.allclose(X[e, 4 + k], 0.5 * (X[e, a] + X[e, b]))
    T = M.LocalMesh2D(2, 2, 1.0, 1.0, P=2).elem_node_xyz
    for e in range(T.shape[0]):
        for k, (a, b) in enumerate(EDGES2):
            assert np.allclose(T[e, 3 + k], 0.5 * (T[e, a] + T[e, b]))


@pytest.mark.parametrize("P", [3, 4])
def test_lagrange_edge_order(P):
    X = M.LocalMesh(2, 2, 2, 1.0, 1.0, 1.0, P=P).elem_node_xyz
    for e in range(X.shape[0]):
        for k, (a, b) in enumerate(EDGES3):
            for j in range(1, P):
                want = ((P - j) * X[e, a] + j * X[e, b]) / P
                assert np.allclose(X[e, 4 + k * (P - 1) + (j - 1)], want)


def test_pvd_and_names(tmp_path):
    assert vtu.frame_name("frame_", 7) == "frame_00007.vtu"
    path = str(tmp_path / "run.pvd")
    vtu.write_pvd(path, ["frame_00000.vtu", "frame_00001.vtu"], [0.0, 0.25])
    txt = open(path).read()
    assert 'timestep="0.25"' in txt and 'file="frame_00001.vtu"' in txt and txt.count("<DataSet") == 2
    with pytest.raises(ValueError):
        vtu.write_pvd(path, ["a.vtu"], [0.0, 1.0])
