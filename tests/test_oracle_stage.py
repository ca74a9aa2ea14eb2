"""Pins the oracle's stage, step and limiter (oracle/stage.c) with what the
mathematics and the reference's own tests fix:

* constant-state preservation for every physics and order
  (test/euler_2d_gpu_test.mojo, test/euler_3d_test.mojo,
  test/sw_2d_gpu_test.mojo, test/maxwell_2d_gpu_test.mojo,
  test/mhd_glm_2d_gpu_test.mojo)
* conservation of the weighted integral on periodic meshes
* the stage combine a*q_a + b*q_b + c*dt*L(q_in) and the SSPRK3 step as the
  cubic Taylor polynomial of the linear advection operator
  (test/ssprk3_test.mojo coefficients; amplification 1+z+z^2/2+z^3/6)
* convergence of the DG derivative of smooth data at rate >= P
* closed-form source terms (gravity, Maxwell J/M) and GLM psi damping
* analytic solutions: Gaussian return, translated vortex, PEC cavity mode,
  Sod plateaus at the Toro values of test/sod_exact_riemann_test.mojo:16-19
* the BJ limiter: passthrough, spike clamp with the far field untouched
  (test/limiter_2d_gpu_test_p3.mojo), cell-mean conservation and a firing
  check (test/limiter_3d_test.mojo), and the scaling form
  q_lim - mean = theta (q - mean) with one theta in [0, 1] per element.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from mojoxm_amd import problems

G = 1.4


def euler_prm(flux=0, efix=False, g=(0.0, 0.0, 0.0)):
    return [G, 1e-6, 1e-6, flux, float(efix), *g]


def const_state(nc, K, NP, vals):
    q = np.empty((K, NP, nc))
    q[...] = np.asarray(vals, float)
    return q


EUL5 = [1.0, 0.5, 0.3, -0.2, 1.0 / (G - 1) + 0.5 * (0.25 + 0.09 + 0.04)]

CASES_3D = [
    ("adv", O.PHYS_ADVECTION, [0.7, -0.4, 0.3], [1.3]),
    *[(f"euler_f{f}_fix{e}", O.PHYS_EULER, euler_prm(f, e), EUL5)
      for f, e in [(0, 0), (1, 0), (2, 0), (3, 0), (1, 1), (3, 1)]],
]
CASES_2D = [
    ("adv", O.PHYS_ADVECTION, [0.7, -0.4, 0.0], [1.3]),
    ("euler_hllec", O.PHYS_EULER, euler_prm(3), [1.0, 0.5, 0.3, 0.0, 1.0 / (G - 1) + 0.5 * 0.34]),
    ("euler_roe_fix", O.PHYS_EULER, euler_prm(1, True), [1.0, 0.5, 0.3, 0.0, 1.0 / (G - 1) + 0.5 * 0.34]),
    ("euler2d_rus", O.PHYS_EULER2D, [G, 1e-6, 1e-6, 0.0], [1.0, 0.5, 0.3, 1.0 / (G - 1) + 0.5 * 0.34]),
    ("euler2d_hllc", O.PHYS_EULER2D, [G, 1e-6, 1e-6, 1.0], [1.0, 0.5, 0.3, 1.0 / (G - 1) + 0.5 * 0.34]),
    ("sw_rus", O.PHYS_SW2D, [9.81, 1e-6, 0.0], [2.0, 0.8, 0.4]),
    ("sw_hll", O.PHYS_SW2D, [9.81, 1e-6, 1.0], [2.0, 0.8, 0.4]),
    ("glm", O.PHYS_MHDGLM2D, [5 / 3, 1e-6, 1e-6, 1.2, 0.0], [1.0, 0.3, -0.2, 0.8, 0.5, 2.0, 0.05]),
    ("maxwell", O.PHYS_MAXWELL2D, [1.0], [0.1, -0.2, 0.3, 0.4, -0.5, 0.6]),
    ("adv2d", O.PHYS_ADVECTION2D, [0.7, -0.4], [1.3]),
]


def _nc(phys):
    return O.lib().orc_phys_nc(phys)


@pytest.mark.parametrize("P", [1, 2, 3, 4])
@pytest.mark.parametrize("name,phys,prm,vals", CASES_3D, ids=[c[0] for c in CASES_3D])
def test_constant_state_3d(name, phys, prm, vals, P):
    m = O.Mesh(3, P, (2, 2, 2), (1.0, 1.2, 0.9))
    q = const_state(_nc(phys), m.K, m.NP, vals)
    out = m.step(phys, prm, q, 1e-3, 2)
    assert np.max(np.abs(out - q)) <= 1e-13 * np.max(np.abs(q))


@pytest.mark.parametrize("P", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("name,phys,prm,vals", CASES_2D, ids=[c[0] for c in CASES_2D])
def test_constant_state_2d(name, phys, prm, vals, P):
    m = O.Mesh(2, P, (4, 3), (1.0, 0.8))
    q = const_state(_nc(phys), m.K, m.NP, vals)
    out = m.step(phys, prm, q, 1e-3, 2)
    assert np.max(np.abs(out - q)) <= 1e-13 * np.max(np.abs(q))


def weighted_integral(m, q):
    """sum_e |e| sum_i w_i q[e, i, :]; every element of these meshes has the
    same size, so the element measure is a common factor."""
    return np.einsum("i,eic->c", m.ref["w"], q)


@pytest.mark.parametrize("dim,name,phys,prm,vals",
                         [(3, *c) for c in CASES_3D[:3]] + [(2, *c) for c in CASES_2D],
                         ids=[f"{d}d_{c[0]}" for d, c in [(3, c) for c in CASES_3D[:3]] + [(2, c) for c in CASES_2D]])
def test_periodic_conservation(dim, name, phys, prm, vals):
    rng = np.random.default_rng(5)
    m = O.Mesh(dim, 2, (3, 3, 2) if dim == 3 else (5, 4), (1.0, 1.0, 1.0) if dim == 3 else (1.0, 1.0))
    q = const_state(_nc(phys), m.K, m.NP, vals)
    q *= 1.0 + 0.05 * rng.uniform(-1, 1, q.shape)
    before = weighted_integral(m, q)
    out = m.step(phys, prm, q, 2e-3, 3)
    after = weighted_integral(m, out)
    assert np.all(np.abs(after - before) <= 1e-13 * np.maximum(1.0, np.abs(before)))


def operator(m, phys, prm, q):
    """L(q) through one stage with a = b = 0, c = dt = 1."""
    return m.stage(phys, prm, q, q, q, 0.0, 0.0, 1.0, 1.0)


@pytest.mark.parametrize("dim", [2, 3])
def test_stage_combine_and_ssprk3_polynomial(dim):
    rng = np.random.default_rng(7)
    phys, prm = O.PHYS_ADVECTION, [0.9, -0.5, 0.4 if dim == 3 else 0.0]
    m = O.Mesh(dim, 2, (2, 2, 2) if dim == 3 else (4, 3), (1.0,) * dim)
    q = rng.normal(size=(m.K, m.NP, 1))
    qa, qb = rng.normal(size=q.shape), rng.normal(size=q.shape)
    Lq = operator(m, phys, prm, q)
    a, b, c, dt = 0.3, 0.45, 0.7, 0.013
    out = m.stage(phys, prm, q, qa, qb, a, b, c, dt)
    assert np.allclose(out, a * qa + b * qb + c * dt * Lq, rtol=0, atol=1e-13)
    # one SSPRK3 step of a linear operator is its cubic Taylor polynomial
    dt = 0.01
    L2q = operator(m, phys, prm, Lq)
    L3q = operator(m, phys, prm, L2q)
    expect = q + dt * Lq + dt ** 2 / 2 * L2q + dt ** 3 / 6 * L3q
    got = m.step(phys, prm, q, dt, 1)
    assert np.max(np.abs(got - expect)) <= 1e-13 * np.max(np.abs(q))


@pytest.mark.parametrize("dim,P", [(2, 2), (2, 3), (2, 4), (3, 2), (3, 3)])
def test_derivative_converges(dim, P):
    """L(q) -> -v . grad q for smooth periodic q, at rate >= P."""
    v = np.array([0.8, -0.6, 0.5][:dim])
    prm = [*v, 0.0][:3] if dim == 3 else [v[0], v[1], 0.0]
    errs = []
    ns = (4, 8)
    for n in ns:
        m = O.Mesh(dim, P, (n,) * dim, (1.0,) * dim)
        x = m.elem_node_xyz
        k = 2 * np.pi
        arg = k * x.sum(-1)
        q = np.sin(arg)[..., None]
        exact = -(v.sum() * k * np.cos(arg))[..., None]
        Lq = operator(m, O.PHYS_ADVECTION, prm, q)
        errs.append(np.sqrt(np.mean((Lq - exact) ** 2)) / np.sqrt(np.mean(exact ** 2)))
    rate = math.log(errs[0] / errs[1], 2)
    assert rate >= P - 0.3, (errs, rate)


def test_gravity_source_closed_form():
    """Uniform state under gravity: rho const, m(t) = m0 + rho g t,
    E(t) = E0 + m0.g t + rho |g|^2 t^2 / 2 -- quadratic in t, so three
    SSPRK3 stages reproduce it exactly."""
    g = np.array([0.3, -0.7, 0.2])
    for dim in (2, 3):
        m = O.Mesh(dim, 2, (2, 2, 2) if dim == 3 else (3, 3), (1.0,) * dim)
        gg = g if dim == 3 else np.array([g[0], g[1], 0.0])
        q0 = np.array([1.2, 0.3, -0.1, 0.2 if dim == 3 else 0.0, 3.0])
        q = const_state(5, m.K, m.NP, q0)
        dt, n = 0.01, 4
        out = m.step(O.PHYS_EULER, euler_prm(0, False, gg), q, dt, n)
        t = dt * n
        expect = q0.copy()
        expect[1:4] += q0[0] * gg * t
        expect[4] += (q0[1:4] @ gg) * t + 0.5 * q0[0] * (gg @ gg) * t * t
        assert np.max(np.abs(out - expect)) <= 1e-13
    # module Euler2D gravity
    m = O.Mesh(2, 2, (3, 3), (1.0, 1.0))
    q0 = np.array([1.2, 0.3, -0.1, 3.0])
    q = const_state(4, m.K, m.NP, q0)
    out = m.step(O.PHYS_EULER2D, [G, 1e-6, 1e-6, 1.0, g[0], g[1]], q, 0.01, 4)
    t = 0.04
    expect = q0.copy()
    expect[1:3] += q0[0] * g[:2] * t
    expect[3] += (q0[1:3] @ g[:2]) * t + 0.5 * q0[0] * (g[:2] @ g[:2]) * t * t
    assert np.max(np.abs(out - expect)) <= 1e-13


def test_maxwell_uniform_sources():
    """dE/dt = -c^2 J, dB/dt = -M on a uniform field (src/local_mesh_2d_gpu_maxwell.mojo:133-142)."""
    c, J, M = 1.5, np.array([0.2, -0.1, 0.3]), np.array([-0.4, 0.05, 0.1])
    m = O.Mesh(2, 3, (3, 3), (1.0, 1.0))
    q0 = np.array([0.1, -0.2, 0.3, 0.4, -0.5, 0.6])
    out = m.step(O.PHYS_MAXWELL2D, [c, *J, *M], const_state(6, m.K, m.NP, q0), 0.01, 5)
    expect = q0 + 0.05 * np.concatenate([-c * c * J, -M])
    assert np.max(np.abs(out - expect)) <= 1e-14


def test_glm_psi_damping():
    """psi *= exp(-alpha_d dt) once per step, after stage 3; a uniform B has
    zero divergence, so damping is the only change."""
    ad, dt, n = 2.5, 0.01, 6
    m = O.Mesh(2, 2, (3, 3), (1.0, 1.0))
    q0 = np.array([1.0, 0.3, -0.2, 0.8, 0.5, 2.0, 0.05])
    out = m.step(O.PHYS_MHDGLM2D, [5 / 3, 1e-6, 1e-6, 1.2, ad], const_state(7, m.K, m.NP, q0), dt, n)
    assert np.max(np.abs(out[..., :6] - q0[:6])) <= 1e-14
    assert np.max(np.abs(out[..., 6] - q0[6] * math.exp(-ad * dt) ** n)) <= 1e-15


def test_gaussian_return_3d():
    """The Gaussian pulse returns to its initial condition after one period
    (v = (1,1,1), unit cube); 8^3 cubes, P = 2."""
    pb = problems.advection_gaussian_3d(N=8, P=2)
    m = O.Mesh(3, 2, pb.n, pb.L)
    q0 = pb.ic(m.elem_node_xyz)
    out = m.step(O.PHYS_ADVECTION, pb.physics.params(), q0, pb.dt, pb.nsteps)
    err = np.linalg.norm(out - q0) / np.linalg.norm(q0)
    assert 1e-3 < err < 0.12, err


def test_vortex_translates():
    pb = problems.euler_vortex_2d(N=16, P=3, t_final=0.5)
    m = O.Mesh(2, 3, pb.n, pb.L)
    q0 = pb.ic(m.elem_node_xyz)
    out = m.step(O.PHYS_EULER, pb.physics.params(), q0, pb.dt, pb.nsteps)
    exact = pb.exact(m.elem_node_xyz, pb.t_final)
    err = np.linalg.norm(out - exact) / np.linalg.norm(exact)
    assert err < 1e-3, err


def test_maxwell_pec_cavity_mode():
    """TM(1,1) mode of the PEC unit square: Ez = sin(pi x) sin(pi y) cos(w t),
    Bx = -(pi/w) sin(pi x) cos(pi y) sin(w t), By = (pi/w) cos(pi x) sin(pi y) sin(w t)."""
    m = O.Mesh(2, 3, (8, 8), (1.0, 1.0), bc=[1, 1, 1, 1])
    x, y = m.elem_node_xyz[..., 0], m.elem_node_xyz[..., 1]
    w = math.pi * math.sqrt(2.0)

    def mode(t):
        q = np.zeros((m.K, m.NP, 6))
        q[..., 2] = np.sin(np.pi * x) * np.sin(np.pi * y) * math.cos(w * t)
        q[..., 3] = -(np.pi / w) * np.sin(np.pi * x) * np.cos(np.pi * y) * math.sin(w * t)
        q[..., 4] = (np.pi / w) * np.cos(np.pi * x) * np.sin(np.pi * y) * math.sin(w * t)
        return q

    dt, n = 0.005, 60
    out = m.step(O.PHYS_MAXWELL2D, [1.0], mode(0.0), dt, n)
    exact = mode(dt * n)
    err = np.linalg.norm(out - exact) / np.linalg.norm(exact)
    assert err < 1e-3, err
    # the PEC walls keep the tangential E and normal B at rest on the boundary
    assert np.max(np.abs(out[..., 0:2])) < 1e-3


def test_sod_plateaus_match_toro():
    """Sod tube on [0,1] (outflow ends, slip walls across), trait Euler HLLE
    with the BJ limiter after every stage, P = 1; at t = 0.2 the star-region
    plateaus match Toro's values p* = 0.30313, rho*L = 0.42632,
    rho*R = 0.26557 (test/sod_exact_riemann_test.mojo:16-19)."""
    nx = 200
    m = O.Mesh(2, 1, (nx, 1), (1.0, 1.0 / nx), bc=[2, 2, 1, 1])
    x = m.elem_node_xyz[..., 0]
    left = x < 0.5
    q = np.zeros((m.K, m.NP, 5))
    q[..., 0] = np.where(left, 1.0, 0.125)
    q[..., 4] = np.where(left, 1.0, 0.1) / (G - 1)
    dt = 0.2 / 800
    out = m.step(O.PHYS_EULER, euler_prm(2), q, dt, 800, limiter=True, eps=0.1)
    xc = x.mean(axis=1)
    rho = out[..., 0].mean(axis=1)
    u = out[..., 1].mean(axis=1) / rho
    p = (G - 1) * (out[..., 4].mean(axis=1) - 0.5 * rho * u * u)

    def plateau(v, a, b):
        sel = (xc > a) & (xc < b)
        return v[sel].mean()

    assert abs(plateau(rho, 0.52, 0.66) - 0.42632) < 0.01
    assert abs(plateau(rho, 0.72, 0.82) - 0.26557) < 0.01
    assert abs(plateau(p, 0.55, 0.82) - 0.30313) < 0.005
    assert abs(plateau(u, 0.55, 0.82) - 0.92745) < 0.02
    assert np.all(np.isfinite(out))


# ----------------------------------------------------------------- limiter --
def test_limiter_2d_passthrough_and_spike():
    m = O.Mesh(2, 3, (6, 4), (1.0, 1.0))
    vals = [1.0, 0.2, 0.1, 2.52]
    q = const_state(4, m.K, m.NP, vals)
    out = m.limiter(4, q, 0.1)
    assert np.max(np.abs(out - q)) <= 1e-14
    q[0, 0, 0] = 5.0
    out = m.limiter(4, q, 0.1)
    assert np.all(np.isfinite(out))
    assert out[..., 0].max() < 2.5
    far = np.arange(m.K) >= m.K // 2
    assert np.max(np.abs(out[far] - q[far])) <= 1e-14
    check_scaling_form(m, q, out, per_component=False)


@pytest.mark.parametrize("P", [2, 3])
def test_limiter_3d_conserves_means(P):
    m = O.Mesh(3, P, (3, 3, 3), (1.0, 1.0, 1.0))
    rng = np.random.default_rng(11)
    q = const_state(5, m.K, m.NP, EUL5)
    out = m.limiter(5, q, 0.0)
    assert np.max(np.abs(out - q)) <= 1e-14
    q = q + 0.4 * rng.uniform(-1, 1, q.shape)
    out = m.limiter(5, q, 0.0)
    w = m.ref["w"]
    assert np.max(np.abs(np.einsum("i,eic->ec", w, out) - np.einsum("i,eic->ec", w, q))) <= 1e-13
    assert np.max(np.abs(out - q)) > 0.01
    check_scaling_form(m, q, out, per_component=False)


def check_scaling_form(m, q, out, per_component):
    w = m.ref["w"]
    mean = np.einsum("i,eic->ec", w, q)[:, None, :]
    dq, dl = q - mean, out - mean
    for e in range(m.K):
        d = dq[e].ravel()
        if np.max(np.abs(d)) < 1e-12:
            continue
        theta = float(d @ dl[e].ravel() / (d @ d))
        assert -1e-12 <= theta <= 1 + 1e-12
        assert np.max(np.abs(dl[e] - theta * dq[e])) <= 1e-12
