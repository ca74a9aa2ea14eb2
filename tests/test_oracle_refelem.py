"""Pins the oracle's reference-element operators (oracle/refelem.c) and the
library's host builder (mxm_refelem_*) against exact rational goldens,
closed forms and the reference's own known-answer tests."""
import os

import numpy as np
import pytest

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "refelem_golden.npz"))


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("P", [1, 2, 3, 4, 5])
def test_oracle_matches_exact_rational_3d(oracle, P):
    r = oracle.refelem3d(P)
    t = f"3d_p{P}"
    assert _rel(r["D"], GOLD["D_" + t]) < 1e-13
    assert _rel(r["Lift"], GOLD["Lift_" + t]) < 1e-13
    assert np.abs(r["w"] - GOLD["w_" + t]).max() < 1e-14
    assert (r["f2e"] == GOLD["fmap_" + t]).all()


@pytest.mark.parametrize("P", [1, 2, 3, 4, 5])
def test_oracle_matches_exact_rational_2d(oracle, P):
    r = oracle.refelem2d(P)
    t = f"2d_p{P}"
    assert _rel(r["D"], GOLD["D_" + t]) < 1e-13
    assert _rel(r["Lift"], GOLD["Lift_" + t]) < 1e-13
    assert np.abs(r["w"] - GOLD["w_" + t]).max() < 1e-14
    assert (r["e2e"] == GOLD["fmap_" + t]).all()


def test_p2_tet_closed_forms(oracle):
    # test/reference_element_test.mojo:133-172: -1/20 at vertices, 1/5 at edges
    w = oracle.refelem3d(2)["w"]
    assert np.allclose(w[:4], -1.0 / 20.0, atol=1e-15)
    assert np.allclose(w[4:], 1.0 / 5.0, atol=1e-15)


def test_p2_triangle_closed_forms(oracle):
    # test/reference_element_2d_test.mojo:111-137: 0 at vertices, 1/3 at edges
    w = oracle.refelem2d(2)["w"]
    assert np.allclose(w[:3], 0.0, atol=1e-15)
    assert np.allclose(w[3:], 1.0 / 3.0, atol=1e-15)


def test_p2_face_table_matches_hand_coded(oracle):
    # src/reference.mojo:699-752 (ref_face_node): the hand-unrolled P=2 table
    hand = np.array([[1, 2, 3, 5, 9, 8], [0, 2, 3, 6, 9, 7], [0, 1, 3, 4, 8, 7], [0, 1, 2, 4, 5, 6]])
    assert (oracle.refelem3d(2)["f2e"] == hand).all()


def test_mat_inv_known_answers(oracle):
    # test/reference_element_test.mojo:175-250
    assert np.allclose(oracle.mat_inv(np.array([[1.0, 2.0], [3.0, 4.0]])), [[-2.0, 1.0], [1.5, -0.5]], atol=1e-12)
    assert np.allclose(oracle.mat_inv(np.eye(3)), np.eye(3), atol=1e-12)
    assert np.allclose(oracle.mat_inv(np.array([[0.0, 2.0], [3.0, 4.0]])),
                       [[-2.0 / 3.0, 1.0 / 3.0], [0.5, 0.0]], atol=1e-12)
    with pytest.raises(ZeroDivisionError):
        oracle.mat_inv(np.zeros((2, 2)))


@pytest.mark.parametrize("P", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("dim", [2, 3])
def test_weak_derivative_identities(oracle, P, dim):
    """sum_i (M D_k)[i, j] = int d(sum_i phi_i)/dr_k phi_j = 0, and D_k applied
    to the nodal values of the linear field r_k gives M^-1 int phi_i (1)'s
    weak form: sum_j D_k[i, j] x_j^(k) = M^-1 int dphi_i/dr_k r_k."""
    r = oracle.refelem3d(P) if dim == 3 else oracle.refelem2d(P)
    D, w = r["D"], r["w"]
    # Row sums of D_k on a constant field: D_k 1 = M^-1 int dphi_i/dr_k = M^-1 (surface term)
    # Weighted column identity: w^T D_k = 0 because sum_i w_i phi_i integrates to
    # the mean and int d(1)/dr_k phi_j = 0.
    for k in range(dim):
        assert np.abs(w @ D[k]).max() < 1e-11 * max(1.0, np.abs(D[k]).max())


def test_library_host_refelem_matches_goldens():
    import mojoxm_amd as M
    for P in range(1, 6):
        r3, r2 = M.ReferenceElement(P), M.ReferenceElement2D(P)
        t3, t2 = f"3d_p{P}", f"2d_p{P}"
        tol = 1e-15 if P <= 3 else 1e-12
        assert _rel(r3.D_ref, GOLD["D_" + t3]) < tol and _rel(r3.Lift_ref, GOLD["Lift_" + t3]) < tol
        assert _rel(r2.D_ref, GOLD["D_" + t2]) < tol and _rel(r2.Lift_ref, GOLD["Lift_" + t2]) < tol
        assert np.abs(r3.node_weights - GOLD["w_" + t3]).max() < 1e-15
        assert np.abs(r2.node_weights - GOLD["w_" + t2]).max() < 1e-15
        assert (r3.face_to_elem == GOLD["fmap_" + t3]).all() and (r2.edge_to_elem == GOLD["fmap_" + t2]).all()
