"""Pins the oracle's approximate Riemann solvers and its limiter function with
closed forms that the mathematics fixes, derived here from first principles
rather than copied from the formulas being checked.

* HLLC family (trait Euler HLLEC, src/euler.mojo:415-485; the Euler2D module's
  HLLC, src/local_mesh_2d_gpu_euler.mojo:466-545).  The two star states are
  rebuilt from the Rankine-Hugoniot conditions across the outer waves, with
  the normal velocity S* and the pressure p* common to both sides:
      mass        rho* (S_K - S*) = rho_K (S_K - u_K)
      tangential  the tangential velocity is continuous across S_K
      momentum    p* = p_K + rho_K (S_K - u_K)(S* - u_K)      (same p* both sides)
      energy      E* (S_K - S*) = E_K (S_K - u_K) + p* S* - p_K u_K
  and the flux is the upwind one, F_K + S_K (U*_K - U_K), on the side of the
  fan that holds x/t = 0.  A slip in the star energy (a flipped sign of
  p_K / (rho_K (S_K - u_K))) or in S* breaks it; the contact tests cannot see
  the energy term because S* - u_K vanishes there.
* HLL family (trait Euler HLLE without the entropy fix, the SW module's HLL).
  With one intermediate state U_hll, the flux satisfies the jump condition
  across both bounding waves: F - F_L = S_L (U_hll - U_L) and
  F_R - F = S_R (U_R - U_hll).  U_hll is taken from the first, the second is
  checked.
* Wave-speed estimates: Einfeldt's (Roe-averaged) for the trait solvers,
  Davis' (min / max of u_n -+ a over both sides) for the 2D modules.
* Rusanov dissipation speed of the 2D modules, read off the mass component:
  |u| + a (Euler2D), |u| + sqrt(g h) (SW), max(|u| + c_f, c_h) (GLM-MHD) with
  the textbook fast speed along x: B along x gives max(a, c_A), B along y
  gives sqrt(a^2 + c_A^2).
* Venkatakrishnan's limiter function (V. Venkatakrishnan, J. Comput. Phys.
  118 (1995) 120-130), in the reference's form alpha(d, D) =
  (D^2 + 2 D d + eps^2) / (D^2 + 2 d^2 + D d + eps^2) with d the node's
  deviation from the cell mean and D the allowed one
  (src/local_mesh_2d_gpu_limiter.mojo:112-133, src/solver.mojo:483-500).
  Closed values: ratio D/d = 1 gives 3/4, 2 gives 1, 1/2 gives 5/11, 0 gives
  0 (Barth-Jespersen flattening); eps^2 = D^2 = d^2 gives 4/5.  A single
  element is set up with hand-chosen neighbour means, and the limited state
  must be mean + theta (q - mean) with theta the smallest of these values.
  2D limits with component 0 only; 3D with every component.
"""
import numpy as np
import pytest

from oracle import oracle as O

G = 1.4
RMIN = PMIN = 1e-8


def close(a, b, tol=1e-12):
    scale = max(1.0, float(np.max(np.abs(b))))
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) <= tol * scale


# ------------------------------------------------------------ Euler, any dim --
def prim(U, g=G):
    """rho, velocity vector, pressure of a conserved Euler state
    [rho, m (dim), E]."""
    rho = U[0]
    u = U[1:-1] / rho
    p = (g - 1.0) * (U[-1] - 0.5 * rho * u @ u)
    return rho, u, p


def phys_flux(U, n, g=G):
    rho, u, p = prim(U, g)
    un = u @ n
    return np.array([rho * un, *(rho * u * un + p * n), (U[-1] + p) * un])


def star_state(U, S, Ss, n, g=G):
    """The HLLC star state on the side of U from the Rankine-Hugoniot
    conditions across the wave of speed S, with normal velocity Ss."""
    rho, u, p = prim(U, g)
    un = u @ n
    rs = rho * (S - un) / (S - Ss)
    ps = p + rho * (S - un) * (Ss - un)
    us = Ss * n + (u - un * n)
    Es = (U[-1] * (S - un) + ps * Ss - p * un) / (S - Ss)
    return np.array([rs, *(rs * us), Es]), ps


def contact_speed(UL, UR, SL, SR, n, g=G):
    """S* from p*_L = p*_R (the momentum jump across each outer wave)."""
    rl, ul, pl = prim(UL, g)
    rr, ur, pr = prim(UR, g)
    a, b = ul @ n, ur @ n
    return (pr - pl + rl * a * (SL - a) - rr * b * (SR - b)) / (rl * (SL - a) - rr * (SR - b))


def hllc_expected(UL, UR, SL, SR, n, g=G):
    FL, FR = phys_flux(UL, n, g), phys_flux(UR, n, g)
    if SL >= 0.0:
        return FL
    if SR <= 0.0:
        return FR
    Ss = contact_speed(UL, UR, SL, SR, n, g)
    UsL, psL = star_state(UL, SL, Ss, n, g)
    UsR, psR = star_state(UR, SR, Ss, n, g)
    assert abs(psL - psR) <= 1e-12 * max(1.0, abs(psL))
    return FL + SL * (UsL - UL) if Ss >= 0.0 else FR + SR * (UsR - UR)


def einfeldt(UL, UR, n, g=G):
    rl, ul, pl = prim(UL, g)
    rr, ur, pr = prim(UR, g)
    al, ar = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    sl, sr = np.sqrt(rl), np.sqrt(rr)
    ut = (sl * ul + sr * ur) / (sl + sr)
    Ht = (sl * (UL[-1] + pl) / rl + sr * (UR[-1] + pr) / rr) / (sl + sr)
    at = np.sqrt((g - 1.0) * (Ht - 0.5 * ut @ ut))
    return min(ul @ n - al, ut @ n - at), max(ur @ n + ar, ut @ n + at)


def davis(UL, UR, n, g=G):
    rl, ul, pl = prim(UL, g)
    rr, ur, pr = prim(UR, g)
    al, ar = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    return min(ul @ n - al, ur @ n - ar), max(ul @ n + al, ur @ n + ar)


def state(rho, u, p, g=G):
    u = np.asarray(u, float)
    return np.array([rho, *(rho * u), p / (g - 1.0) + 0.5 * rho * u @ u])


def rand_pair(rng, dim, regime):
    """A left/right pair: 'sub' keeps both sides subsonic along e_x,
    'right' / 'left' move both far to one side of the fan."""
    u0 = {"sub": 0.0, "right": 4.0, "left": -4.0}[regime]
    out = []
    for _ in range(2):
        u = rng.uniform(-0.6, 0.6, dim)
        u[0] += u0
        out.append(state(rng.uniform(0.3, 3.0), u, rng.uniform(0.3, 3.0)))
    return out


def rand_frame(rng, dim):
    q, r = np.linalg.qr(rng.normal(size=(dim, dim)))
    return q @ np.diag(np.sign(np.diag(r)))


def rotate(U, R):
    return np.array([U[0], *(R @ U[1:-1]), U[-1]])


def trait_flux(U_l, U_r, n, flux):
    """Trait Euler numerical flux (no entropy fix) of two 3D states."""
    f, _ = O.numerical_flux(O.PHYS_EULER, [G, RMIN, PMIN, flux, 0.0, 0.0, 0.0, 0.0] + [0.0] * 5, U_l, U_r, n)
    return f


@pytest.mark.parametrize("regime", ["sub", "right", "left"])
def test_trait_hllec_star_states_from_rankine_hugoniot(regime):
    """HLLEC without the entropy fix equals the HLLC flux of the
    Rankine-Hugoniot star states with Einfeldt's speeds, in a random frame."""
    rng = np.random.default_rng(1000 + len(regime))
    nsign = [0, 0]
    for _ in range(300):
        UL, UR = rand_pair(rng, 3, regime)
        R = rand_frame(rng, 3)
        n = R[:, 0]
        UL, UR = rotate(UL, R), rotate(UR, R)
        SL, SR = einfeldt(UL, UR, n)
        want = hllc_expected(UL, UR, SL, SR, n)
        got = trait_flux(UL, UR, n, 3)
        assert close(got, want, 1e-11), (got, want)
        if regime == "sub" and SL < 0 < SR:
            nsign[int(contact_speed(UL, UR, SL, SR, n) >= 0)] += 1
    if regime == "sub":
        assert min(nsign) > 30  # both star states exercised


@pytest.mark.parametrize("regime", ["sub", "right", "left"])
def test_trait_hlle_jump_conditions(regime):
    """HLLE without the fix: one intermediate state satisfying the jump
    condition across both Einfeldt waves."""
    rng = np.random.default_rng(1100 + len(regime))
    for _ in range(300):
        UL, UR = rand_pair(rng, 3, regime)
        R = rand_frame(rng, 3)
        n = R[:, 0]
        UL, UR = rotate(UL, R), rotate(UR, R)
        SL, SR = einfeldt(UL, UR, n)
        FL, FR = phys_flux(UL, n), phys_flux(UR, n)
        F = trait_flux(UL, UR, n, 2)
        if SL >= 0:
            assert close(F, FL)
        elif SR <= 0:
            assert close(F, FR)
        else:
            Uh = UL + (F - FL) / SL
            assert close(FR - F, SR * (UR - Uh), 1e-11)


@pytest.mark.parametrize("regime", ["sub", "right", "left"])
def test_module_hllc_star_states_from_rankine_hugoniot(regime):
    """The Euler2D module's HLLC with Davis' speeds."""
    rng = np.random.default_rng(1200 + len(regime))
    prm = [G, RMIN, PMIN, 1.0]
    nsign = [0, 0]
    for _ in range(300):
        UL, UR = rand_pair(rng, 2, regime)
        R = rand_frame(rng, 2)
        n = R[:, 0]
        UL, UR = rotate(UL, R), rotate(UR, R)
        SL, SR = davis(UL, UR, n)
        want = hllc_expected(UL, UR, SL, SR, n)
        got = O.face_flux_2d(O.PHYS_EULER2D, prm, UL, UR, 0, n)
        assert close(got, want, 1e-11), (got, want)
        if regime == "sub" and SL < 0 < SR:
            nsign[int(contact_speed(UL, UR, SL, SR, n) >= 0)] += 1
    if regime == "sub":
        assert min(nsign) > 30


@pytest.mark.parametrize("regime", ["sub", "right", "left"])
def test_module_sw_hll_jump_conditions(regime):
    g = 9.81
    rng = np.random.default_rng(1300 + len(regime))
    u0 = {"sub": 0.0, "right": 12.0, "left": -12.0}[regime]

    def fl(U, n):
        h, m = U[0], U[1:]
        u = m / h
        return np.array([m @ n, *(m * (u @ n) + 0.5 * g * h * h * n)])

    for _ in range(300):
        n = rand_frame(rng, 2)[:, 0]
        Us = []
        for _s in range(2):
            h = rng.uniform(0.3, 3.0)
            u = rng.uniform(-1, 1, 2) + u0 * n
            Us.append(np.array([h, *(h * u)]))
        UL, UR = Us
        cl, cr = np.sqrt(g * UL[0]), np.sqrt(g * UR[0])
        unl, unr = UL[1:] @ n / UL[0], UR[1:] @ n / UR[0]
        SL, SR = min(unl - cl, unr - cr), max(unl + cl, unr + cr)
        F = O.face_flux_2d(O.PHYS_SW2D, [g, 1e-8, 1.0], UL, UR, 0, n)
        FL, FR = fl(UL, n), fl(UR, n)
        if SL >= 0:
            assert close(F, FL)
        elif SR <= 0:
            assert close(F, FR)
        else:
            Uh = UL + (F - FL) / SL
            assert close(FR - F, SR * (UR - Uh), 1e-11)


def rusanov_alpha(phys, prm, ql, qr, n):
    """Dissipation speed of a Rusanov flux, from its mass component."""
    f = O.face_flux_2d(phys, prm, ql, qr, 0, n)
    c = 0.5 * (O.face_flux_2d(phys, prm, ql, ql, 0, n) + O.face_flux_2d(phys, prm, qr, qr, 0, n))
    return -2.0 * (f[0] - c[0]) / (qr[0] - ql[0])


def test_module_rusanov_speeds_closed_forms():
    n = np.array([0.6, 0.8])
    # Euler2D: |u| + a, the larger side
    ql, qr = state(1.0, [0.3, -0.4], 1.0), state(2.0, [0.0, 0.1], 1.5)
    want = max(0.5 + np.sqrt(G), 0.1 + np.sqrt(G * 1.5 / 2.0))
    assert abs(rusanov_alpha(O.PHYS_EULER2D, [G, RMIN, PMIN, 0.0], ql, qr, n) - want) < 1e-12
    # SW: |u| + sqrt(g h)
    g = 9.81
    ql, qr = np.array([1.0, 0.3, -0.4]), np.array([2.0, 0.0, 0.2])
    want = max(0.5 + np.sqrt(g), 0.1 + np.sqrt(2 * g))
    assert abs(rusanov_alpha(O.PHYS_SW2D, [g, 1e-8, 0.0], ql, qr, n) - want) < 1e-12
    # GLM-MHD: fast speed along x; B along x -> max(a, c_A), B along y -> sqrt(a^2 + c_A^2)
    gm = 5.0 / 3.0

    def mhd(rho, u, p, B, psi=0.0):
        u, B = np.asarray(u, float), np.asarray(B, float)
        return np.array([rho, *(rho * u), *B, p / (gm - 1) + 0.5 * rho * u @ u + 0.5 * B @ B, psi])

    for B, fast in (([2.0, 0.0], lambda a, ca: max(a, ca)), ([0.0, 2.0], lambda a, ca: np.hypot(a, ca)),
                    ([0.3, 0.0], lambda a, ca: max(a, ca))):
        rho, p = 1.3, 0.9
        a, ca = np.sqrt(gm * p / rho), np.linalg.norm(B) / np.sqrt(rho)
        ql = mhd(rho, [0.0, 0.0], p, B)
        qr = mhd(rho * 0.5, [0.0, 0.0], p * 0.5, B)  # same a, c_A larger on the right
        ca_r = np.linalg.norm(B) / np.sqrt(rho * 0.5)
        want = max(fast(a, ca), fast(a, ca_r))
        for ch in (0.1, want + 0.5):
            got = rusanov_alpha(O.PHYS_MHDGLM2D, [gm, RMIN, PMIN, ch, 0.0], ql, qr, n)
            assert abs(got - max(want, ch)) < 1e-12, (B, ch, got, want)


# ----------------------------------------------------------------- limiter --
def venkat_case(dim, P, nc, setup):
    """Builds a periodic mesh, picks element e*, gives each element a
    constant state (the neighbours of e* the means setup() asks for, every
    other element the mean m) and e* the nodal deviations setup() gives;
    returns the mesh, e*, the input and the limited state."""
    n = (4, 4) if dim == 2 else (2, 2, 2)
    m = O.Mesh(dim, P, n, (1.0,) * dim)
    e = 5
    nbr = []
    for f in m.elem_faces[e]:
        a, b = m.face_elem[f]
        nbr.append(int(b if a == e else a))
    assert len(set(nbr)) == len(nbr) and e not in nbr
    mean, nbr_means, dev, eps = setup(m)
    q = np.tile(mean, (m.K, m.NP, 1)).astype(float)
    for k, v in zip(nbr, nbr_means):
        q[k] = v
    assert np.all(np.abs(m.ref["w"] @ dev) < 1e-15)  # the deviations keep e*'s mean
    q[e] = mean + dev
    return m, e, q, m.limiter(nc, q, eps)


def limited_as(q_in, q_out, e, mean, theta):
    # elements with a constant state stay put (up to the rounding of their
    # own weighted mean, which the limiter may flatten them to)
    others = np.arange(q_in.shape[0]) != e
    assert close(q_out[others], q_in[others], 1e-14)
    assert close(q_out[e], mean + theta * (q_in[e] - mean), 1e-14)


@pytest.mark.parametrize("ratio_pos,ratio_neg,eps_over_d,theta", [
    (1.0, 2.0, 0.0, 0.75),      # D = d: 3/4; the other side at D = 2 d: 1
    (0.5, 2.0, 0.0, 5.0 / 11),  # D = d / 2: 5/11
    (2.0, 2.0, 0.0, 1.0),       # D = 2 d on both sides: no limiting
    (0.0, 2.0, 0.0, 0.0),       # D = 0: Barth-Jespersen flattening to the mean
    (1.0, 2.0, 1.0, 0.8),       # eps^2 = D^2 = d^2: 4/5
])
def test_venkatakrishnan_closed_forms_2d(ratio_pos, ratio_neg, eps_over_d, theta):
    """2D, P = 1 (three nodes, weights 1/3): node 0 sits d above the mean,
    nodes 1 and 2 sit d/2 below it.  One neighbour's mean lies
    D_pos = ratio_pos * d above e*'s, another's D_neg = ratio_neg * d/2
    below, so the ratios D/|delta| are ratio_pos and ratio_neg.  Only
    component 0 sets theta (component 1's large deviation does not); every
    component is scaled by it."""
    d = 0.2
    mean = np.array([1.0, 3.0, -2.0, 7.0])

    def setup(m):
        assert np.allclose(m.ref["w"], 1.0 / 3.0)
        dev = np.zeros((3, 4))
        dev[:, 0] = [d, -d / 2, -d / 2]
        dev[:, 1] = [-40 * d, 20 * d, 20 * d]  # ignored by the 2D theta
        dev[:, 3] = [0.0, d, -d]
        # the negative nodes deviate by d/2; ratio_neg is D_neg / (d/2)
        hi, lo = mean.copy(), mean.copy()
        hi[0] += ratio_pos * d
        lo[0] -= ratio_neg * d / 2
        return mean, [hi, lo, mean], dev, eps_over_d * d

    m, e, q, out = venkat_case(2, 1, 4, setup)
    limited_as(q, out, e, mean, theta)


def test_venkatakrishnan_3d_all_components():
    """3D, P = 1 (four nodes, weights 1/4), every component counts.
    Component 2 deviates by d on every node with D = d on both sides
    (ratio 1: 3/4); component 4 deviates by +-d on two nodes with
    D_pos = d/2 (ratio 1/2: 5/11) and D_neg = 2 d (ratio 2: 1).  theta is
    the smallest, 5/11; without component 4's deviation it is 3/4."""
    d = 0.1
    mean = np.array([1.0, -0.5, 2.0, 0.0, 4.0])

    def setup(m):
        assert np.allclose(m.ref["w"], 0.25)
        dev = np.zeros((4, 5))
        dev[:, 2] = [d, -d, d, -d]      # |delta| = d on every node
        dev[:, 4] = [d, -d, 0.0, 0.0]   # |delta| = d on two nodes
        hi, lo, mid = mean.copy(), mean.copy(), mean.copy()
        hi[2] += d          # comp 2: D_pos = D_neg = d -> 3/4
        lo[2] -= d
        hi[4] += d / 2      # comp 4: D_pos = d/2 -> 5/11; D_neg = 2 d -> 1
        lo[4] -= 2 * d
        return mean, [hi, lo, mid, mid], dev, 0.0

    m, e, q, out = venkat_case(3, 1, 5, setup)
    limited_as(q, out, e, mean, 5.0 / 11)
    # without component 4's deviation, component 2 sets 3/4
    q2 = q.copy()
    q2[e, :, 4] = mean[4]
    out2 = m.limiter(5, q2, 0.0)
    limited_as(q2, out2, e, mean, 0.75)
