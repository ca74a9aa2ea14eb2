"""Mesh connectivity: invariants that pin the oracle's restatement of
src/local_mesh.mojo / src/local_mesh_2d.mojo (after
test/local_mesh_2d_test.mojo:36-106 and their 3D analogues), and bit-exact
agreement of the library's independently derived builder with it."""
import os

import numpy as np
import pytest

import mojoxm_amd as M

BC3 = [[0] * 6, [1, 1, 0, 0, 0, 0], [2, 0, 3, 1, 0, 2], [0, 1, 0, 0, 2, 0], [1, 2, 3, 1, 2, 3]]
BC2 = [[0] * 4, [1, 1, 0, 0], [2, 0, 3, 1], [0, 1, 0, 0], [3, 2, 1, 1]]
INT_TABLES = ("elem_faces", "elem_face_side", "elem_canon_to_ref", "face_elem", "face_elem_node", "face_bc_type")
GEO_TABLES = ("elem_node_xyz", "elem_invJ", "elem_inv_vol", "face_normal", "face_area")


def _wrap_diff(a, b, L):
    d = a - b
    return d - np.round(d / L) * L


def check_invariants(m, ndim, n, L, bc):
    K, NF, NFP = m.K, m.NF, m.NFP
    ne_expect = (6 if ndim == 3 else 2) * int(np.prod(n))
    assert K == ne_expect
    # every element's face references resolve back to the element
    for e in range(K):
        for lf in range(NF):
            f, s = m.elem_faces[e, lf], m.elem_face_side[e, lf]
            if m.face_bc_type[f] == 0 or s == 0:
                assert m.face_elem[f, s] == e
    # face node coordinates coincide on both sides (periodic faces)
    xyz = m.elem_node_xyz
    for f in range(m.Nf):
        if m.face_bc_type[f] != 0:
            continue
        e0, e1 = m.face_elem[f]
        p0 = xyz[e0, m.face_elem_node[f, 0]]
        p1 = xyz[e1, m.face_elem_node[f, 1]]
        for d in range(ndim):
            assert np.abs(_wrap_diff(p0[:, d], p1[:, d], L[d])).max() < 1e-12
    # canon_to_ref is consistent with the reference face map
    fmap = m.ref["f2e"] if ndim == 3 else m.ref["e2e"]
    for e in range(K):
        for lf in range(NF):
            f, s = m.elem_faces[e, lf], m.elem_face_side[e, lf]
            for q in range(NFP):
                r = m.elem_canon_to_ref[e, lf, q]
                assert m.face_elem_node[f, s, q] == fmap[lf, r]
    # closed surfaces: sum_f sign * area * n = 0 per element (with the side
    # sign, since normals point out of side 0)
    tot = np.zeros((K, ndim))
    for lf in range(NF):
        f = m.elem_faces[:, lf]
        sgn = np.where(m.elem_face_side[:, lf] == 0, 1.0, -1.0)[:, None]
        tot += sgn * m.face_area[f][:, None] * m.face_normal[f]
    assert np.abs(tot).max() < 1e-12
    # normals point away from side 0: toward the side-1 centroid (periodic)
    cen = xyz.mean(axis=1)
    for f in range(m.Nf):
        if m.face_bc_type[f]:
            continue
        e0, e1 = m.face_elem[f]
        d = np.array([_wrap_diff(cen[e1, k], cen[e0, k], L[k]) for k in range(ndim)])
        assert d @ m.face_normal[f] > 0
    # positive Jacobians
    assert (m.elem_inv_vol > 0).all()


@pytest.mark.parametrize("P", [1, 2, 3])
@pytest.mark.parametrize("bc", BC3)
def test_oracle_mesh3d_invariants(oracle, P, bc):
    n, L = (3, 2, 2), (1.0, 0.7, 1.3)
    m = oracle.Mesh(3, P, n, L, bc)
    check_invariants(m, 3, n, L, bc)
    base = 12 * int(np.prod(n))
    extra = (2 * n[1] * n[2] if bc[0] else 0) + (2 * n[0] * n[2] if bc[2] else 0) + (2 * n[0] * n[1] if bc[4] else 0)
    assert m.Nf == base + extra
    assert (m.face_bc_type[base:] != 0).all()


@pytest.mark.parametrize("P", [1, 2, 3, 4])
@pytest.mark.parametrize("bc", BC2)
def test_oracle_mesh2d_invariants(oracle, P, bc):
    n, L = (4, 3), (2.0, 1.5)
    m = oracle.Mesh(2, P, n, L, bc)
    check_invariants(m, 2, n, L, bc)
    assert m.Nf == 3 * 12 + (3 if bc[0] else 0) + (4 if bc[2] else 0)


def test_oracle_mesh3d_p2_literal_values(oracle):
    """Spot values fixed by the reference tables: face 0 of cube 0 (ft 0)
    joins tet 0 (side 0) and tet 5 (side 1); the normal of ft 0 is
    (0, 1, -1)/sqrt 2 on the unit cube, area sqrt(2)/2."""
    m = oracle.Mesh(3, 2, (2, 2, 2), (1.0, 1.0, 1.0))
    assert tuple(m.face_elem[0]) == (0, 5)
    h = 0.5
    assert np.allclose(m.face_area[0], 0.5 * np.sqrt(2.0) * h * h)
    assert np.allclose(np.abs(m.face_normal[0]), [0.0, 1 / np.sqrt(2), 1 / np.sqrt(2)])
    # +x faces of cube 0 (ft 6, 7) point along +x and join cube 0 to cube 1
    assert np.allclose(m.face_normal[6], [1, 0, 0]) and np.allclose(m.face_normal[7], [1, 0, 0])
    assert m.face_elem[6, 1] // 6 == 1 and m.face_elem[7, 1] // 6 == 1


@pytest.mark.parametrize("P", [1, 2, 3])
@pytest.mark.parametrize("bc", BC3)
def test_library_mesh3d_bit_exact(oracle, P, bc):
    n, L = (3, 4, 2), (1.0, 2.0, 3.0)
    pm = M.LocalMesh(*n, *L, bcs=M.BoundaryConditions(*bc), P=P)
    om = oracle.Mesh(3, P, n, L, bc)
    for k in INT_TABLES:
        assert (getattr(pm, k) == getattr(om, k)).all(), k
    for k in GEO_TABLES:
        assert np.abs(getattr(pm, k) - getattr(om, k)).max() <= 4e-15 * max(1.0, np.abs(getattr(om, k)).max()), k


@pytest.mark.parametrize("P", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("bc", BC2)
def test_library_mesh2d_bit_exact(oracle, P, bc):
    n, L = (5, 3), (2.0, 1.5)
    pm = M.LocalMesh2D(*n, *L, bcs=M.BoundaryConditions2D(*bc), P=P)
    om = oracle.Mesh(2, P, n, L, bc)
    for k in INT_TABLES:
        assert (getattr(pm, k) == getattr(om, k)).all(), k
    for k in GEO_TABLES:
        assert np.abs(getattr(pm, k) - getattr(om, k)).max() <= 4e-15 * max(1.0, np.abs(getattr(om, k)).max()), k


def test_library_mesh3d_bit_exact_p4_p5(oracle):
    for P in (4, 5):
        pm = M.LocalMesh(2, 2, 3, 1.0, 1.0, 1.0, bcs=M.BoundaryConditions(1, 2, 0, 0, 3, 1), P=P)
        om = oracle.Mesh(3, P, (2, 2, 3), (1.0, 1.0, 1.0), [1, 2, 0, 0, 3, 1])
        for k in INT_TABLES:
            assert (getattr(pm, k) == getattr(om, k)).all(), (P, k)


def test_config_sizes(oracle):
    """C1 and C3 element/face counts (16^3 and 32^3 cubes x 6 tets)."""
    m = M.LocalMesh(16, 16, 16, P=2)
    assert (m.num_elements, m.num_faces) == (24576, 49152)
    m2 = M.LocalMesh2D(128, 128, 10.0, 10.0, P=3)
    assert (m2.num_elements, m2.num_faces) == (32768, 49152)


@pytest.mark.parametrize("n", [(1, 2, 2), (2, 1, 2), (2, 2, 1), (1, 1, 1), (3, 1, 2)])
@pytest.mark.parametrize("bc", [[0] * 6, [1, 2, 0, 0, 0, 0], [0, 0, 0, 0, 1, 1]])
def test_library_mesh3d_one_cell_per_period(oracle, n, bc):
    """A periodic axis of one cube: a face's two sides are the same cube's
    opposite faces.  The library's tables equal the oracle's.  (The
    centroid-direction invariant is not defined here: across a one-cube
    period the two centroids are half a period apart either way.)"""
    L = (1.0, 2.0, 1.5)
    for P in (1, 2, 3):
        pm = M.LocalMesh(*n, *L, bcs=M.BoundaryConditions(*bc), P=P)
        om = oracle.Mesh(3, P, n, L, bc)
        for k in INT_TABLES:
            assert (getattr(pm, k) == getattr(om, k)).all(), (P, k)


@pytest.mark.parametrize("n", [(1, 3), (4, 1), (1, 1)])
@pytest.mark.parametrize("bc", [[0] * 4, [1, 2, 0, 0], [0, 0, 3, 1]])
def test_library_mesh2d_one_cell_per_period(oracle, n, bc):
    L = (2.0, 1.5)
    for P in (1, 2, 3):
        pm = M.LocalMesh2D(*n, *L, bcs=M.BoundaryConditions2D(*bc), P=P)
        om = oracle.Mesh(2, P, n, L, bc)
        for k in INT_TABLES:
            assert (getattr(pm, k) == getattr(om, k)).all(), (P, k)


# ------------------------------------------- fixtures from the literal tables --
GOLDEN_MESH = os.path.join(os.path.dirname(__file__), "golden", "kuhn_mesh_golden.npz")


@pytest.mark.parametrize("P", [1, 2, 3])
def test_kuhn_mesh_matches_literal_table_fixtures(oracle, P):
    """The oracle's and the library's periodic 2x2x2 meshes equal the
    connectivity that tests/golden/make_mesh_golden.py builds geometrically
    from the reference's literal Kuhn tables (kuhn_vertex, tet_face_info,
    face_type_corner, corner_d*; src/local_mesh.mojo:64-292), read from its
    source as text."""
    g = np.load(GOLDEN_MESH)
    want = {k: g[f"{k}_p{P}"] for k in ("elem_faces", "elem_face_side", "face_elem", "face_elem_node",
                                        "elem_canon_to_ref", "face_normal", "face_area")}
    meshes = {"oracle": oracle.Mesh(3, P, (2, 2, 2), (1.0, 1.0, 1.0)), "library": M.LocalMesh(2, 2, 2, P=P)}
    for who, m in meshes.items():
        for k, v in want.items():
            got = getattr(m, k)
            if v.dtype.kind == "i":
                assert np.array_equal(got, v), (who, k)
            else:
                assert np.abs(got - v).max() < 1e-15, (who, k)
        assert np.array_equal(m.face_bc_type, np.zeros(m.Nf, np.int32)), who
