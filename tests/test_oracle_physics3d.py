"""The oracle's other trait physics (oracle/physics3d.c: shallow water,
Maxwell, ideal MHD with GLM, five-moment two-fluid), pinned by properties
the reference formulas must have: consistency F*(q, q, n) = F(q).n,
conservation F*(qL, qR, n) = -F*(qR, qL, -n), rotational invariance of the
vector fields, reduction to the Euler Rusanov flux (MHD with B = psi = 0;
the two-fluid species with the fields and charges off), the wall ghosts'
zero mass flux and the PEC fixed point, the sources in closed form, and at
the stage level constant states, the reductions again, and a Maxwell plane
wave converging under refinement."""
import math

import numpy as np
import pytest

from oracle import oracle as O

from mojoxm_amd import physics as ph

PHYS = {
    "sw": ph.ShallowWater(g=9.81, h_min=1e-6),
    "maxwell": ph.Maxwell(c=1.3, J=(0.1, -0.2, 0.3), M=(0.05, 0.0, -0.1)),
    "mhd": ph.IdealMHD(gamma=5.0 / 3.0, c_h=1.7, alpha_d=2.0),
    "twofluid": ph.FiveMomentTwoFluid(c_light=2.0, c_h=1.5, alpha_d=0.5),
}


def rand_state(name, rng):
    if name == "sw":
        return np.array([rng.uniform(0.8, 1.5), *rng.uniform(-0.4, 0.4, 2)])
    if name == "maxwell":
        return rng.uniform(-1, 1, 6)
    if name == "mhd":
        rho = rng.uniform(0.8, 1.3)
        u = rng.uniform(-0.3, 0.3, 3)
        B = rng.uniform(-0.6, 0.6, 3)
        p = rng.uniform(0.6, 1.2)
        E = p / (2.0 / 3.0) + 0.5 * rho * u @ u + 0.5 * B @ B
        return np.array([rho, *(rho * u), E, *B, rng.uniform(-0.1, 0.1)])
    q = np.zeros(17)
    for s0 in (0, 5):
        rho = rng.uniform(0.8, 1.3)
        u = rng.uniform(-0.3, 0.3, 3)
        p = rng.uniform(0.6, 1.2)
        q[s0:s0 + 5] = [rho, *(rho * u), p / (2.0 / 3.0) + 0.5 * rho * u @ u]
    q[10:17] = rng.uniform(-0.5, 0.5, 7)
    return q


def unit(rng):
    n = rng.standard_normal(3)
    return n / np.linalg.norm(n)


@pytest.mark.parametrize("name", list(PHYS))
def test_consistency_and_conservation(name):
    P = PHYS[name]
    rng = np.random.default_rng(1)
    for _ in range(20):
        q, q2 = rand_state(name, rng), rand_state(name, rng)
        n = unit(rng)
        if name == "sw":
            n[2] = 0.0
            n /= np.linalg.norm(n)
        F = O.internal_flux(P.phys_id, P.params(), q)
        f, _ = O.numerical_flux(P.phys_id, P.params(), q, q, n)
        assert np.allclose(f, n @ F, rtol=1e-13, atol=1e-13)
        a, _ = O.numerical_flux(P.phys_id, P.params(), q, q2, n)
        b, _ = O.numerical_flux(P.phys_id, P.params(), q2, q, -n)
        assert np.allclose(a, -b, rtol=1e-13, atol=1e-13)


def rot(axis, ang):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


def rotate_state(name, q, R):
    q = q.copy()
    vec = {"maxwell": [0, 3], "mhd": [1, 5], "twofluid": [1, 6, 10, 13]}[name]
    for o in vec:
        q[o:o + 3] = R @ q[o:o + 3]
    return q


@pytest.mark.parametrize("name", ["maxwell", "mhd", "twofluid"])
def test_rotational_invariance(name):
    P = PHYS[name]
    rng = np.random.default_rng(2)
    for _ in range(10):
        ql, qr, n = rand_state(name, rng), rand_state(name, rng), unit(rng)
        R = rot(rng.standard_normal(3), rng.uniform(0, 2 * math.pi))
        f, _ = O.numerical_flux(P.phys_id, P.params(), ql, qr, n)
        g, _ = O.numerical_flux(P.phys_id, P.params(), rotate_state(name, ql, R), rotate_state(name, qr, R), R @ n)
        assert np.allclose(g, rotate_state(name, f, R), rtol=1e-12, atol=1e-12)


def test_mhd_without_field_is_euler_rusanov():
    mhd = ph.IdealMHD(gamma=1.4, c_h=0.0)
    eu = ph.Euler(gamma=1.4, flux_type=ph.FLUX_RUSANOV)
    rng = np.random.default_rng(3)
    for _ in range(10):
        ql, qr = rand_state("mhd", rng), rand_state("mhd", rng)
        for q in (ql, qr):
            q[4] -= 0.5 * q[5:8] @ q[5:8]
            q[5:9] = 0.0
        n = unit(rng)
        f, _ = O.numerical_flux(mhd.phys_id, mhd.params(), ql, qr, n)
        e, _ = O.numerical_flux(eu.phys_id, eu.params(), ql[:5], qr[:5], n)
        assert np.allclose(f[:5], e, rtol=1e-12, atol=1e-12) and np.all(f[5:] == 0.0)


def test_twofluid_species_are_euler_rusanov():
    tf = ph.FiveMomentTwoFluid(gamma_e=1.4, gamma_i=1.4, c_light=0.0, c_h=0.0)
    eu = ph.Euler(gamma=1.4, flux_type=ph.FLUX_RUSANOV)
    rng = np.random.default_rng(4)
    for _ in range(10):
        ql, qr = rand_state("twofluid", rng), rand_state("twofluid", rng)
        for q in (ql, qr):
            q[5:10] = q[0:5]  # identical species: one shared Rusanov speed
            q[10:] = 0.0
        n = unit(rng)
        f, _ = O.numerical_flux(tf.phys_id, tf.params(), ql, qr, n)
        e, _ = O.numerical_flux(eu.phys_id, eu.params(), ql[:5], qr[:5], n)
        assert np.allclose(f[:5], e, rtol=1e-12, atol=1e-12)
        assert np.allclose(f[5:10], e, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name", ["sw", "mhd", "twofluid"])
def test_wall_zero_mass_flux(name):
    P = PHYS[name]
    rng = np.random.default_rng(5)
    for _ in range(10):
        q, n = rand_state(name, rng), unit(rng)
        if name == "sw":
            n[2] = 0.0
            n /= np.linalg.norm(n)
        f, _ = O.boundary_flux(P.phys_id, P.params(), q, 1, n)
        assert abs(f[0]) < 1e-13
        if name == "twofluid":
            assert abs(f[5]) < 1e-13


@pytest.mark.parametrize("name", ["maxwell", "twofluid"])
def test_pec_fixed_point(name):
    """A state with E along n, B tangential (and fluids at rest) is its own
    PEC wall ghost, so the wall flux is the physical flux; reversing E
    breaks that."""
    P = PHYS[name]
    rng = np.random.default_rng(6)
    for _ in range(10):
        n = unit(rng)
        t = np.cross(n, unit(rng))
        E, B = rng.uniform(0.2, 1) * n, rng.uniform(0.2, 1) * t
        if name == "maxwell":
            q = np.concatenate([E, B])
        else:
            q = np.zeros(17)
            q[0], q[4], q[5], q[9] = 1.1, 2.0, 0.9, 1.5
            q[10:13], q[13:16], q[16] = E, B, 0.1
        f, _ = O.boundary_flux(P.phys_id, P.params(), q, 1, n)
        assert np.allclose(f, n @ O.internal_flux(P.phys_id, P.params(), q), atol=1e-13)
        q2 = q.copy()
        o = 0 if name == "maxwell" else 10
        q2[o:o + 3] = -E + 0.3 * t  # a tangential E is flipped by the ghost
        f2, _ = O.boundary_flux(P.phys_id, P.params(), q2, 1, n)
        assert not np.allclose(f2, n @ O.internal_flux(P.phys_id, P.params(), q2), atol=1e-6)


def test_mhd_speeds_closed_form():
    """Rusanov speed = max(|u.n| + c_f over both sides, c_h), read back from
    the dissipation, with the textbook fast speeds: c_f = max(a, c_A) for B
    along n, c_f = sqrt(a^2 + c_A^2) for B across n, a^2 = gamma p / rho."""
    g = 5.0 / 3.0
    n = np.array([0.0, 0.6, 0.8])
    t = np.array([1.0, 0.0, 0.0])
    def st(rho, p, B, un=0.0):
        u = un * n
        return np.array([rho, *(rho * u), p / (g - 1) + 0.5 * rho * u @ u + 0.5 * B @ B, *B, 0.0])
    for ch in (0.0, 4.0):
        mhd = ph.IdealMHD(gamma=g, c_h=ch)
        for B, cf in ((0.5 * n, lambda a, ca: max(a, ca)), (2.0 * n, lambda a, ca: max(a, ca)),
                      (0.7 * t, lambda a, ca: math.hypot(a, ca))):
            ql, qr = st(1.0, 1.0, B, 0.1), st(0.8, 0.7, B, -0.2)
            want = max(0.1 + cf(math.sqrt(g), np.linalg.norm(B)),
                       0.2 + cf(math.sqrt(g * 0.7 / 0.8), np.linalg.norm(B) / math.sqrt(0.8)), ch)
            f, sp = O.numerical_flux(mhd.phys_id, mhd.params(), ql, qr, n)
            assert sp == pytest.approx(want, rel=1e-13)
            F = n @ O.internal_flux(mhd.phys_id, mhd.params(), ql) + n @ O.internal_flux(mhd.phys_id, mhd.params(), qr)
            assert np.allclose(f, 0.5 * F - 0.5 * want * (qr - ql), rtol=1e-12, atol=1e-13)


def test_inflow_and_outflow_ghosts():
    for name, P in PHYS.items():
        rng = np.random.default_rng(7)
        q, n = rand_state(name, rng), unit(rng)
        if name == "sw":
            n[2] = 0.0
            n /= np.linalg.norm(n)
        f, _ = O.boundary_flux(P.phys_id, P.params(), q, 2, n)  # outflow: ghost = q
        assert np.allclose(f, n @ O.internal_flux(P.phys_id, P.params(), q), rtol=1e-13, atol=1e-13)
        # inflow: the ghost is the stored state
        g = rand_state(name, rng)
        P2 = type(P)(**{**P.__dict__, "inflow": tuple(g)})
        f2, _ = O.boundary_flux(P2.phys_id, P2.params(), q, 3, n)
        want, _ = O.numerical_flux(P2.phys_id, P2.params(), q, g, n)
        assert np.allclose(f2, want, rtol=1e-14, atol=1e-14)


def test_sources_closed_form():
    mw = ph.Maxwell(c=2.0, J=(1.0, -2.0, 0.5), M=(0.25, 0.0, -1.0))
    assert np.allclose(O.source(mw.phys_id, mw.params(), np.zeros(6)), [-4.0, 8.0, -2.0, -0.25, 0.0, 1.0])
    mhd = ph.IdealMHD(alpha_d=3.0)
    q = np.zeros(9)
    q[0], q[8] = 1.0, 0.5
    assert np.allclose(O.source(mhd.phys_id, mhd.params(), q), [0] * 8 + [-1.5])
    tf = ph.FiveMomentTwoFluid(q_e=-1.0, m_e=0.5, q_i=1.0, m_i=2.0, eps0=0.5, alpha_d=2.0)
    q = np.zeros(17)
    q[0], q[1] = 2.0, 2.0          # electrons: rho 2, u = x
    q[5] = 1.0                     # ions at rest
    q[12], q[15], q[16] = 1.0, 1.0, 0.25   # E = z, B = z, psi
    s = O.source(tf.phys_id, tf.params(), q)
    # electrons: q/m = -2, rho = 2: force -4 (E + u x B) = -4 (z - y); power -4 (u . E) = 0
    assert np.allclose(s[0:5], [0, 0, 4, -4, 0])
    # ions at rest: q/m = 0.5, rho = 1: force 0.5 E = 0.5 z; power 0
    assert np.allclose(s[5:10], [0, 0, 0, 0.5, 0])
    # J = (-2)(2)(x) = -4 x; dE/dt = -J / eps0 = 8 x; psi damping -0.5
    assert np.allclose(s[10:17], [8, 0, 0, 0, 0, 0, -0.5])
    q[7] = 1.0                     # ions moving along y: u x B = y x z = x
    s = O.source(tf.phys_id, tf.params(), q)
    # ions: 0.5 (E + x) = (0.5, 0, 0.5); power 0.5 (y . z) = 0
    assert np.allclose(s[5:10], [0, 0.5, 0, 0.5, 0])
    # J = -4 x + 0.5 y; dE/dt = -J / eps0 = (8, -1, 0)
    assert np.allclose(s[10:17], [8, -1, 0, 0, 0, 0, -0.5])


def constant_state(name):
    if name == "sw":
        return np.array([1.2, 0.3, -0.2])
    if name == "maxwell":
        return np.array([0.3, -0.1, 0.2, 0.5, 0.1, -0.4])
    if name == "mhd":
        return np.array([1.1, 0.2, -0.1, 0.05, 3.0, 0.4, -0.3, 0.2, 0.0])
    q = np.zeros(17)
    q[0:5] = [1.1, 0.0, 0.0, 0.0, 2.0]
    q[5:10] = [0.9, 0.0, 0.0, 0.0, 1.5]
    q[13:16] = [0.2, 0.1, -0.3]
    return q


@pytest.mark.parametrize("name", list(PHYS))
def test_constant_state_is_steady(name):
    P = PHYS[name]
    if name == "maxwell":
        P = ph.Maxwell(c=1.3)
    m = O.Mesh(3, 2, (2, 2, 2), (1.0, 1.0, 1.0))
    q = np.broadcast_to(constant_state(name), (m.K, m.NP, P.NUM_COMPONENTS)).copy()
    out = m.step(P.phys_id, P.params(), q, 1e-3, 2)
    assert np.abs(out - q).max() < 1e-13


def test_mhd_stage_reduces_to_euler():
    rng = np.random.default_rng(8)
    m = O.Mesh(3, 2, (2, 2, 2), (1.0, 1.0, 1.0))
    q9 = np.zeros((m.K, m.NP, 9))
    rho = rng.uniform(0.9, 1.1, (m.K, m.NP))
    u = rng.uniform(-0.2, 0.2, (m.K, m.NP, 3))
    p = rng.uniform(0.9, 1.1, (m.K, m.NP))
    q9[..., 0], q9[..., 1:4] = rho, rho[..., None] * u
    q9[..., 4] = p / 0.4 + 0.5 * rho * (u ** 2).sum(-1)
    mhd = ph.IdealMHD(gamma=1.4, c_h=0.0)
    eu = ph.Euler(gamma=1.4, flux_type=ph.FLUX_RUSANOV)
    a = m.step(mhd.phys_id, mhd.params(), q9, 2e-3, 2)
    b = m.step(eu.phys_id, eu.params(), q9[..., :5], 2e-3, 2)
    assert np.abs(a[..., :5] - b).max() < 1e-12 and np.all(a[..., 5:] == 0.0)


def test_maxwell_plane_wave_converges():
    """E_y = cos(2 pi (x - c t)), B_z = E_y / c on the periodic unit cube."""
    c = 1.0
    mw = ph.Maxwell(c=c)
    errs = []
    for n in (4, 8):
        m = O.Mesh(3, 2, (n, n, n), (1.0, 1.0, 1.0))
        x = m.elem_node_xyz[..., 0]
        def wave(t):
            q = np.zeros((m.K, m.NP, 6))
            q[..., 1] = np.cos(2 * math.pi * (x - c * t))
            q[..., 5] = q[..., 1] / c
            return q
        T, steps = 0.1, 5 * n
        out = m.step(mw.phys_id, mw.params(), wave(0.0), T / steps, steps)
        errs.append(np.linalg.norm(out - wave(T)) / np.linalg.norm(wave(T)))
    assert errs[1] < errs[0] / 5 and errs[1] < 1e-2


def test_twofluid_speed_is_largest_signal():
    """Rusanov speed = max over both species of |u.n| + a_s, c_light and c_h
    (src/two_fluid.mojo:384-386), species at rest with a_s = sqrt(gamma p / rho)."""
    n = np.array([0.0, 0.0, 1.0])
    q = np.zeros(17)
    q[0], q[4] = 1.0, 1.0 / (5.0 / 3.0 - 1)   # electrons: p = 1
    q[5], q[9] = 2.0, 0.5 / (1.4 - 1)         # ions: p = 0.5, rho = 2
    for cl, ch, want in ((5.0, 1.0, 5.0), (0.5, 3.0, 3.0), (0.1, 0.2, math.sqrt(5.0 / 3.0))):
        tf = ph.FiveMomentTwoFluid(gamma_e=5.0 / 3.0, gamma_i=1.4, c_light=cl, c_h=ch)
        _, sp = O.numerical_flux(tf.phys_id, tf.params(), q, q, n)
        assert sp == pytest.approx(want, rel=1e-14)


def test_langmuir_uniform_plasma_exact():
    """examples/two_fluid_langmuir.mojo: a uniform plasma has no flux
    divergence, so the stage is the source ODE; its (E_x, u_e, u_i) follow
    the Langmuir oscillation at w^2 = n0 (1 + m_e / m_i) in closed form
    (SSPRK3's O(dt^3) error is far below the tolerance here)."""
    from mojoxm_amd import problems as pr
    p = pr.langmuir_3d(t_final=0.2)
    m = O.Mesh(3, p.P, p.n, p.L)
    out = m.step(p.physics.phys_id, p.physics.params(), p.ic(m.elem_node_xyz), p.dt, p.nsteps)
    E, ue, ui = p.notes["exact_fields"](p.t_final)
    assert out[..., 10].mean() == pytest.approx(E, rel=1e-10)
    assert (out[..., 1] / out[..., 0]).mean() == pytest.approx(ue, rel=1e-10)
    assert (out[..., 6] / out[..., 5]).mean() == pytest.approx(ui, rel=1e-8)
    assert out[..., 10].std() < 1e-15


def test_maxwell_cavity_with_pec_walls_converges():
    """examples/maxwell_cavity.mojo: the standing mode between PEC walls at
    y = 0, 1 converges at the P = 2 rate (error / 8 per halving of h)."""
    from mojoxm_amd import problems as pr
    errs = []
    for NY in (8, 16):
        p = pr.maxwell_cavity_3d(NY=NY, t_final=0.5)
        m = O.Mesh(3, p.P, p.n, p.L, [0, 0, 1, 1, 0, 0])
        out = m.step(p.physics.phys_id, p.physics.params(), p.ic(m.elem_node_xyz), p.dt, p.nsteps)
        ex = p.exact(m.elem_node_xyz, p.t_final)
        errs.append(np.linalg.norm(out - ex) / np.linalg.norm(ex))
    assert errs[1] < errs[0] / 6 and errs[1] < 2e-4
