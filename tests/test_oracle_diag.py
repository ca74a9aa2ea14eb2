"""The oracle's diagnostics (oracle/diag.c) and nodal differentiation
matrices, pinned by what the mathematics fixes: derivatives of monomials,
integrals of polynomial fields (exact under the mass-weighted nodal rule
for degree <= P), rigid rotation (curl u = 2 Omega), and the Taylor-Green
initial field's KE = pi^3 and enstrophy = 3 pi^3 on the 2 pi cube, reached
under refinement.  The library's own nodal derivative (refelem.cpp, a
different construction) must agree with the oracle's."""
import itertools
import math

import numpy as np
import pytest

from oracle import oracle as O

import mojoxm_amd as M
from mojoxm_amd import problems


def monomials(dim, P):
    for e in itertools.product(range(P + 1), repeat=dim):
        if sum(e) <= P:
            yield e


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("P", [1, 2, 3, 4, 5])
def test_nodal_deriv_exact_on_monomials(dim, P):
    Dn = O.nodal_deriv(dim, P)
    X = (O.refelem3d(P) if dim == 3 else O.refelem2d(P))["node_pos"]
    for e in monomials(dim, P):
        f = np.prod([X[:, d] ** e[d] for d in range(dim)], axis=0)
        for k in range(dim):
            if e[k] == 0:
                want = np.zeros(len(X))
            else:
                ek = list(e)
                ek[k] -= 1
                want = e[k] * np.prod([X[:, d] ** ek[d] for d in range(dim)], axis=0)
            assert np.abs(Dn[k] @ f - want).max() < 1e-11 * max(1, P ** 2), (e, k)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("P", [1, 2, 3, 4, 5])
def test_library_nodal_deriv_matches_oracle(dim, P):
    r = M.ReferenceElement(P) if dim == 3 else M.ReferenceElement2D(P)
    assert np.abs(r.D_nodal - O.nodal_deriv(dim, P)).max() < 1e-13


def box_volume(m):
    return float(np.prod(m.L))


@pytest.mark.parametrize("dim,phys,nc", [(2, O.PHYS_EULER, 5), (2, O.PHYS_EULER2D, 4), (2, O.PHYS_MAXWELL2D, 6),
                                         (3, O.PHYS_EULER, 5), (3, O.PHYS_ADVECTION, 1)])
def test_constant_state(dim, phys, nc):
    n, L = ((3, 4), (2.0, 3.0)) if dim == 2 else ((2, 3, 2), (1.0, 2.0, 1.5))
    m = O.Mesh(dim, 2, n, L)
    v = np.linspace(0.7, 1.9, nc)
    v[0] = 1.3
    q = np.broadcast_to(v, (m.K, m.NP, nc)).copy()
    out = m.diagnostics(phys, q)
    V = box_volume(m)
    assert np.allclose(out[:nc], v * V, rtol=1e-13)
    assert np.allclose(out[nc:2 * nc], v * v * V, rtol=1e-13)
    assert np.array_equal(out[2 * nc:3 * nc], np.abs(v))
    if phys in (O.PHYS_EULER, O.PHYS_EULER2D):
        nv = 3 if phys == O.PHYS_EULER else 2
        u = v[1:1 + nv] / v[0]
        assert out[3 * nc] == pytest.approx(0.5 * v[0] * (u @ u) * V, rel=1e-13)
        assert abs(out[3 * nc + 1]) < 1e-20
    else:
        assert np.isnan(out[3 * nc]) and np.isnan(out[3 * nc + 1])


@pytest.mark.parametrize("P", [2, 3])
def test_polynomial_integrals_exact(P):
    """f = 1 + x + 2 y^2 (degree 2 <= P): int f over [0,2]x[0,3] = 48; the
    nodal rule also integrates f^2 exactly when 4 <= P -- not here -- so
    only the square of the linear field g = x + 2 y (degree 2) is exact:
    int g^2 = int x^2 + 4xy + 4y^2 = 8 + 36 + 72 = 116 (for P >= 2)."""
    m = O.Mesh(2, P, (4, 3), (2.0, 3.0))
    x, y = m.elem_node_xyz[..., 0], m.elem_node_xyz[..., 1]
    q = np.zeros((m.K, m.NP, 1))
    q[..., 0] = 1 + x + 2 * y * y
    assert m.diagnostics(O.PHYS_ADVECTION2D, q)[0] == pytest.approx(48.0, rel=1e-13)
    q[..., 0] = x + 2 * y
    assert m.diagnostics(O.PHYS_ADVECTION2D, q)[1] == pytest.approx(116.0, rel=1e-13)


@pytest.mark.parametrize("P", [2, 3])
def test_rigid_rotation(P):
    """u = Omega (-(y - y0), x - x0), rho const: curl u = 2 Omega, so the
    enstrophy is 1/2 rho (2 Omega)^2 |Omega|-area exactly; KE = 1/2 rho
    Omega^2 int r^2 is a quadratic, exact for P >= 2."""
    om, rho = 0.7, 1.5
    for dim in (2, 3):
        n, L = ((3, 4), (2.0, 3.0)) if dim == 2 else ((2, 3, 2), (2.0, 3.0, 1.0))
        m = O.Mesh(dim, P, n, L)
        X = m.elem_node_xyz
        dx, dy = X[..., 0] - 1.0, X[..., 1] - 1.5
        q = np.zeros((m.K, m.NP, 5))
        q[..., 0] = rho
        q[..., 1] = -om * dy * rho
        q[..., 2] = om * dx * rho
        q[..., 4] = 2.5
        out = m.diagnostics(O.PHYS_EULER, q)
        V = box_volume(m)
        ir2 = (L[0] ** 3 / 12 * L[1] + L[1] ** 3 / 12 * L[0]) * (L[2] if dim == 3 else 1.0)
        assert out[15] == pytest.approx(0.5 * rho * om * om * ir2, rel=1e-12)
        assert out[16] == pytest.approx(0.5 * rho * (2 * om) ** 2 * V, rel=1e-12)


def test_taylor_green_converges():
    """KE = pi^3 and enstrophy = 3 pi^3 (rho = 1 initially).  The KE
    integrand is a trigonometric product that the equispaced nodal rule on
    the periodic lattice integrates exactly once N >= 4; the enstrophy
    carries the interpolation error of the derivative and converges under
    h refinement at P = 3."""
    errs = []
    for N in (2, 4, 8):
        m = O.Mesh(3, 3, (N, N, N), (2 * math.pi,) * 3)
        q = problems.taylor_green_ic(m.elem_node_xyz)
        out = m.diagnostics(O.PHYS_EULER, q)
        errs.append((abs(out[15] / math.pi ** 3 - 1), abs(out[16] / (3 * math.pi ** 3) - 1)))
    assert errs[0][0] > 0.1 and errs[1][0] < 1e-13 and errs[2][0] < 1e-13
    assert errs[2][1] < 1e-2 and errs[1][1] < errs[0][1] / 4 and errs[2][1] < errs[1][1] / 8


def test_mass_weighted_differs_from_unweighted_mean():
    """The reference's unweighted nodal mean (src/diagnostics.mojo:170) is
    not the integral at P >= 2; the mass-weighted rule is."""
    m = O.Mesh(2, 2, (2, 2), (1.0, 1.0))
    x = m.elem_node_xyz[..., 0]
    q = (x * x)[..., None]
    weighted = m.diagnostics(O.PHYS_ADVECTION2D, q)[0]
    unweighted = q.sum() * 1.0 / q.size
    assert weighted == pytest.approx(1.0 / 3.0, rel=1e-13)
    assert abs(unweighted - 1.0 / 3.0) > 1e-3
