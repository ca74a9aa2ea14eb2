"""Initial conditions, analytic solutions and time-step rules of the driver
programs (host FP64), and the benchmark configurations built from them.

  gaussian_ic          examples/advection_gaussian.mojo:82-120
  taylor_green_ic      examples/euler_taylor_green.mojo:86-129
  isentropic_vortex    examples/euler_vortex_2d_gpu.mojo:90-118 (Shu form;
                       ``reference_factor=True`` reproduces the reference
                       drivers' temperature factor (gamma-1) beta^2 /
                       (8 gamma (2 pi)^2), one quarter of Shu's)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from . import physics as ph
from .boundary import BoundaryConditions2D
from .time_integrator import apriori_dt, fixed_steps


def _nearest(d: np.ndarray, L: float) -> np.ndarray:
    d = np.where(d > 0.5 * L, d - L, d)
    return np.where(d < -0.5 * L, d + L, d)


def gaussian_ic(xyz: np.ndarray, L=(1.0, 1.0, 1.0), center=(0.5, 0.5, 0.5), sigma=0.1) -> np.ndarray:
    """exp(-|x - c|^2 / (2 sigma^2)) with the nearest periodic image; [K, NP, 1]."""
    r2 = sum(_nearest(xyz[..., d] - center[d], L[d]) ** 2 for d in range(xyz.shape[-1]))
    return np.exp(-r2 / (2.0 * sigma * sigma))[..., None]


def taylor_green_ic(xyz: np.ndarray, u0=1.0, rho0=1.0, mach=0.3, gamma=1.4) -> np.ndarray:
    """Compressible TGV: u = u0 sin x cos y cos z, v = -u0 cos x sin y cos z,
    p = p0 + rho0 u0^2/16 (cos 2x + cos 2y)(cos 2z + 2), p0 = rho0 u0^2/(gamma Ma^2)."""
    x, y, z = xyz[..., 0], xyz[..., 1], xyz[..., 2]
    p0 = rho0 * u0 * u0 / (gamma * mach * mach)
    u = u0 * np.sin(x) * np.cos(y) * np.cos(z)
    v = -u0 * np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(u)
    rho = np.full_like(u, rho0)
    p = p0 + (rho0 * u0 * u0 / 16.0) * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0)
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return np.stack([rho, rho * u, rho * v, rho * w, E], axis=-1)


def isentropic_vortex(xy: np.ndarray, t=0.0, beta=5.0, gamma=1.4, u0=1.0, v0=1.0, center=(5.0, 5.0),
                      L=(10.0, 10.0), t_inf=1.0, nc=5, reference_factor=False) -> np.ndarray:
    """Shu's isentropic vortex advected by (u0, v0) for time t on a periodic
    box; nc=5 gives (rho, rhou, rhov, rhow=0, E), nc=4 drops rhow."""
    two_pi = 2.0 * math.pi
    factor = (gamma - 1.0) * beta * beta / (8.0 * gamma * (two_pi * two_pi if reference_factor else math.pi ** 2))
    dx = np.mod(xy[..., 0] - center[0] - u0 * t + 0.5 * L[0], L[0]) - 0.5 * L[0]
    dy = np.mod(xy[..., 1] - center[1] - v0 * t + 0.5 * L[1], L[1]) - 0.5 * L[1]
    r2 = dx * dx + dy * dy
    T = t_inf - factor * np.exp(1.0 - r2)
    e_half = np.exp(0.5 * (1.0 - r2))
    u = u0 - (beta / two_pi) * dy * e_half
    v = v0 + (beta / two_pi) * dx * e_half
    rho = T ** (1.0 / (gamma - 1.0))
    p = rho * T
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    cols = [rho, rho * u, rho * v] + ([np.zeros_like(rho)] if nc == 5 else []) + [E]
    return np.stack(cols, axis=-1)


@dataclass
class Problem:
    """One configured run: mesh, order, physics, IC, time steps."""
    name: str
    n: tuple
    L: tuple
    P: int
    physics: object
    ic: Callable[[np.ndarray], np.ndarray]
    dt: float
    nsteps: int
    exact: Optional[Callable[[np.ndarray, float], np.ndarray]] = None
    notes: dict = field(default_factory=dict)
    bcs: Optional[object] = None  # BoundaryConditions / BoundaryConditions2D; None = periodic

    @property
    def ndim(self) -> int:
        return len(self.n)

    @property
    def t_final(self) -> float:
        return self.dt * self.nsteps


def advection_gaussian_3d(N=16, P=2, cfl=0.3, t_final=1.0, sigma=0.1) -> Problem:
    """3D scalar advection of a Gaussian pulse, v = (1,1,1), upwind, periodic
    unit cube, returning to its initial condition at T = 1."""
    phys = ph.Advection(1.0, 1.0, 1.0)
    n, dt = fixed_steps(t_final, apriori_dt(cfl, 1.0 / N, math.sqrt(3.0), P))
    ic = lambda xyz: gaussian_ic(xyz, sigma=sigma)
    return Problem(f"advection_gaussian_3d_n{N}_p{P}", (N, N, N), (1.0, 1.0, 1.0), P, phys, ic, dt, n,
                   exact=lambda xyz, t: gaussian_ic(xyz, sigma=sigma))


def euler_vortex_2d(N=128, P=3, cfl=0.15, t_final=2.0, beta=5.0, flux=ph.FLUX_HLLEC,
                    reference_factor=False) -> Problem:
    """2D isentropic vortex on periodic [0,10]^2 with the trait Euler (NC=5,
    w = 0, nz = 0), background velocity (1,1)."""
    gamma = 1.4
    phys = ph.Euler(gamma=gamma, min_density=1e-6, min_pressure=1e-6, flux_type=flux)
    h = 10.0 / N
    wave = math.sqrt(2.0) + math.sqrt(gamma * 1.0) + beta / (2.0 * math.pi)
    n, dt = fixed_steps(t_final, apriori_dt(cfl, h, wave, P))
    kw = dict(beta=beta, gamma=gamma, reference_factor=reference_factor)
    return Problem(f"euler_vortex_2d_n{N}_p{P}", (N, N), (10.0, 10.0), P, phys,
                   lambda xy: isentropic_vortex(xy, 0.0, **kw), dt, n,
                   exact=lambda xy, t: isentropic_vortex(xy, t, **kw))


def taylor_green_3d(N=32, P=3, cfl=0.2, nsteps=200, mach=0.3, flux=ph.FLUX_RUSANOV) -> Problem:
    """Compressible Taylor-Green vortex on the periodic 2 pi cube."""
    gamma = 1.4
    phys = ph.Euler(gamma=gamma, min_density=1e-6, min_pressure=1e-6, flux_type=flux)
    c0 = math.sqrt(gamma * (1.0 / (gamma * mach * mach)) / 1.0)
    dt = apriori_dt(cfl, 2.0 * math.pi / N, 1.0 + c0, P)
    L = 2.0 * math.pi
    return Problem(f"taylor_green_3d_n{N}_p{P}", (N, N, N), (L, L, L), P, phys,
                   lambda xyz: taylor_green_ic(xyz, mach=mach, gamma=gamma), dt, nsteps)


# ---------------------------------------------------------------------------
# The 3D drivers of the other trait physics (tetrahedra, P = 2, dt = CFL h /
# (wave (2P+1)) as in each driver's choose_dt).
# ---------------------------------------------------------------------------
def maxwell_cavity_3d(NY=32, P=2, cfl=0.2, t_final=2.0, c=1.0) -> Problem:
    """examples/maxwell_cavity.mojo:1-80: the lowest standing mode between
    PEC walls at y = 0 and y = 1, periodic in x and z (NY/2 x NY x 2 cubes,
    z thickness 2/NY): E_x = sin(pi y) cos(w t), B_z = cos(pi y) sin(w t) / c,
    w = c pi.  One period T = 2 / c returns the IC."""
    from .boundary import BoundaryConditions
    n, L = (NY // 2, NY, 2), (1.0, 1.0, 2.0 / NY)
    steps, dt = fixed_steps(t_final, apriori_dt(cfl, 1.0 / NY, c, P))

    def exact(xyz, t):
        y = xyz[..., 1]
        q = np.zeros(xyz.shape[:-1] + (6,))
        q[..., 0] = np.sin(math.pi * y) * math.cos(c * math.pi * t)
        q[..., 5] = np.cos(math.pi * y) * math.sin(c * math.pi * t) / c
        return q
    return Problem(f"maxwell_cavity_3d_ny{NY}_p{P}", n, L, P, ph.Maxwell(c=c), lambda xyz: exact(xyz, 0.0), dt, steps,
                   exact=exact, bcs=BoundaryConditions(0, 0, 1, 1, 0, 0))


def mhd_alfven_3d(NX=32, P=2, cfl=0.2, t_final=1.0, amp=0.1, gamma=5.0 / 3.0, rho0=1.0, B0=1.0, p0=0.1, c_h=1.5,
                  alpha_d=0.5) -> Problem:
    """examples/mhd_alfven.mojo:1-133: a transverse Alfven wave along +x on
    the periodic box NX x 4 x 4 cubes of side 1/NX, B = (B0, 0, 0),
    u_y = A sin(k x - w t), B_y = -A sin(k x - w t), w = k B0 / sqrt(rho0),
    k = 2 pi; exact for the linearised equations (the finite amplitude adds
    an O(A^2) magnetic-pressure disturbance)."""
    n, L = (NX, 4, 4), (1.0, 4.0 / NX, 4.0 / NX)
    cf = math.sqrt(gamma * p0 / rho0 + (B0 * B0 + amp * amp) / rho0)
    steps, dt = fixed_steps(t_final, apriori_dt(cfl, 1.0 / NX, max(cf, c_h), P))
    cA, k = B0 / math.sqrt(rho0), 2.0 * math.pi

    def exact(xyz, t):
        s = amp * np.sin(k * xyz[..., 0] - cA * k * t)
        q = np.zeros(xyz.shape[:-1] + (9,))
        q[..., 0] = rho0
        q[..., 2] = rho0 * s
        q[..., 5] = B0
        q[..., 6] = -s
        q[..., 4] = p0 / (gamma - 1.0) + 0.5 * rho0 * s * s + 0.5 * (B0 * B0 + s * s)
        return q
    phys = ph.IdealMHD(gamma=gamma, min_density=1e-6, min_pressure=1e-6, c_h=c_h, alpha_d=alpha_d)
    return Problem(f"mhd_alfven_3d_nx{NX}_p{P}", n, L, P, phys, lambda xyz: exact(xyz, 0.0), dt, steps, exact=exact)


def langmuir_3d(NX=16, P=2, cfl=0.2, t_final=math.pi, u0=0.01, n0=1.0, m_i=25.0, c=10.0, c_h=12.0, p0=0.01) -> Problem:
    """examples/two_fluid_langmuir.mojo:1-143: a uniform two-fluid plasma
    (q_e = -1, m_e = 1, q_i = 1, eps0 = 1) on the periodic box NX x 2 x 2
    cubes of side 1/NX, electrons started at u_e = (u0, 0, 0).  Uniform in
    space, so the fluxes cancel and the sources drive a Langmuir
    oscillation at w^2 = n0 (q_e^2/m_e + q_i^2/m_i) / eps0:
    E_x = (u0 / w) sin(w t), u_e = u0 - (u0 / w^2)(1 - cos w t),
    u_i = (u0 / (m_i w^2))(1 - cos w t)."""
    gam = 5.0 / 3.0
    n, L = (NX, 2, 2), (1.0, 2.0 / NX, 2.0 / NX)
    steps, dt = fixed_steps(t_final, apriori_dt(cfl, 1.0 / NX, max(c, c_h), P))
    w = math.sqrt(n0 * (1.0 + 1.0 / m_i))
    phys = ph.FiveMomentTwoFluid(gamma_e=gam, gamma_i=gam, q_e=-1.0, m_e=1.0, q_i=1.0, m_i=m_i, eps0=1.0,
                                 c_light=c, c_h=c_h, alpha_d=0.5, min_density=1e-6, min_pressure=1e-8)

    def ic(xyz):
        q = np.zeros(xyz.shape[:-1] + (17,))
        q[..., 0], q[..., 1] = n0, n0 * u0
        q[..., 4] = p0 / (gam - 1.0) + 0.5 * n0 * u0 * u0
        q[..., 5], q[..., 9] = m_i * n0, p0 / (gam - 1.0)
        return q

    def exact_fields(t):
        """(E_x, u_e, u_i) at time t."""
        return (u0 / w * math.sin(w * t), u0 - u0 / (w * w) * (1.0 - math.cos(w * t)),
                u0 / (m_i * w * w) * (1.0 - math.cos(w * t)))
    return Problem(f"langmuir_3d_nx{NX}_p{P}", n, L, P, phys, ic, dt, steps,
                   notes={"omega": w, "exact_fields": exact_fields})


def shallow_water_drop_3d(N=32, P=2, cfl=0.2, t_final=1.0, g=1.0, h0=1.0, amp=0.2, sigma=0.07) -> Problem:
    """examples/shallow_water_drop.mojo:1-97: a Gaussian bump
    h = h0 + a exp(-r^2 / sigma^2) at the centre of the unit square basin
    with slip walls on x and y, periodic in z (N x N x 2 cubes, z thickness
    2/N), at rest.  Total mass is conserved."""
    from .boundary import BoundaryConditions
    n, L = (N, N, 2), (1.0, 1.0, 2.0 / N)
    steps, dt = fixed_steps(t_final, apriori_dt(cfl, 1.0 / N, math.sqrt(g * (h0 + amp)), P))

    def ic(xyz):
        r2 = (xyz[..., 0] - 0.5) ** 2 + (xyz[..., 1] - 0.5) ** 2
        q = np.zeros(xyz.shape[:-1] + (3,))
        q[..., 0] = h0 + amp * np.exp(-r2 / (sigma * sigma))
        return q
    return Problem(f"shallow_water_drop_3d_n{N}_p{P}", n, L, P, ph.ShallowWater(g=g, h_min=1e-4), ic, dt, steps,
                   bcs=BoundaryConditions(1, 1, 1, 1, 0, 0))


# ---------------------------------------------------------------------------
# Batched ensemble: linearly polarised Alfven waves with GLM cleaning.
# ---------------------------------------------------------------------------
# Directions (a, b) of integer wave vectors 2 pi (a, b), gcd 1, 0 <= angle <
# pi/2: the waves are periodic on the unit square.
_ALFVEN_DIRS = [(1, 0), (3, 1), (2, 1), (3, 2), (1, 1), (2, 3), (1, 2), (1, 3)]


def alfven_wave(xy: np.ndarray, t: float, a: int, b: int, amp: float, gamma=5.0 / 3.0, rho0=1.0, p0=0.1,
                B0=1.0) -> np.ndarray:
    """Linear Alfven wave travelling along k = 2 pi (a, b) with the
    background field B0 along k (examples/mhd_alfven_glm_2d_gpu.mojo:98-115
    rotated): u = A sin(phi) t_hat, B = B0 k_hat - A sqrt(rho0) sin(phi) t_hat,
    phi = k . x - |k| c_A t, c_A = B0 / sqrt(rho0); psi = 0.  NC = 7:
    (rho, rho u, rho v, Bx, By, E, psi)."""
    kn = math.hypot(a, b)
    kx, ky = a / kn, b / kn
    tx, ty = -ky, kx
    cA = B0 / math.sqrt(rho0)
    phi = 2.0 * math.pi * (a * xy[..., 0] + b * xy[..., 1]) - 2.0 * math.pi * kn * cA * t
    s = amp * np.sin(phi)
    u, v = s * tx, s * ty
    Bx = B0 * kx - math.sqrt(rho0) * s * tx
    By = B0 * ky - math.sqrt(rho0) * s * ty
    rho = np.full_like(s, rho0)
    E = p0 / (gamma - 1.0) + 0.5 * rho * (u * u + v * v) + 0.5 * (Bx * Bx + By * By)
    return np.stack([rho, rho * u, rho * v, Bx, By, E, np.zeros_like(s)], axis=-1)


@dataclass
class EnsembleMember:
    """One instance of the batched ensemble."""
    index: int
    n: tuple
    L: tuple
    P: int
    physics: object
    dt: float
    nsteps: int
    ic: Callable[[np.ndarray], np.ndarray]
    exact: Callable[[np.ndarray, float], np.ndarray]
    params: dict = field(default_factory=dict)


def alfven_ensemble(count=512, seed=0, P=3, t_final=1.0, cfl=0.15, cr=0.18, n_min=16, n_max=96):
    """The seeded batched ensemble: `count` 2D ideal-MHD + GLM (NC = 7)
    Alfven-wave problems on periodic unit squares, P = 3.  Per instance:
    mesh side n = round(n_min (n_max/n_min)^u), u ~ U[0,1) (log-uniform
    integer in [n_min, n_max], 2 n^2 triangles); wave angle ~ U[0, pi/2),
    taken to the nearest direction of _ALFVEN_DIRS so that the wave is
    periodic; amplitude ~ U[0.01, 0.2].  Background rho = 1, p = 0.1,
    B0 = 1, gamma = 5/3; c_h = the largest signal speed of the initial
    state, alpha_d = c_h / cr; dt = CFL h / (c_h (2P+1)), steps to T."""
    rng = np.random.default_rng(seed)
    gamma, rho0, p0, B0 = 5.0 / 3.0, 1.0, 0.1, 1.0
    dirs = np.array([math.atan2(b, a) for a, b in _ALFVEN_DIRS])
    out = []
    for i in range(count):
        n = int(round(n_min * (n_max / n_min) ** rng.uniform(0.0, 1.0)))
        theta = rng.uniform(0.0, 0.5 * math.pi)
        amp = rng.uniform(0.01, 0.2)
        a, b = _ALFVEN_DIRS[int(np.argmin(np.abs(dirs - theta)))]
        # fast speed bound: |u| + sqrt(cs^2 + |B|^2 / rho)
        c_h = amp + math.sqrt(gamma * p0 / rho0 + (B0 * B0 + amp * amp) / rho0)
        phys = ph.MHDGLM2D(gamma=gamma, c_h=c_h, alpha_d=c_h / cr)
        nsteps, dt = fixed_steps(t_final, apriori_dt(cfl, 1.0 / n, c_h, P))
        kw = dict(a=a, b=b, amp=amp, gamma=gamma, rho0=rho0, p0=p0, B0=B0)
        out.append(EnsembleMember(i, (n, n), (1.0, 1.0), P, phys, dt, nsteps,
                                  lambda xy, kw=kw: alfven_wave(xy, 0.0, **kw),
                                  lambda xy, t, kw=kw: alfven_wave(xy, t, **kw),
                                  dict(theta=theta, direction=(a, b), amplitude=amp, c_h=c_h)))
    return out


# ---------------------------------------------------------------------------
# Serving stream: a seeded Poisson stream of three request types.
# ---------------------------------------------------------------------------
@dataclass
class Request:
    """One solve request of the serving stream."""
    index: int
    kind: str
    arrival: float  # seconds after the stream starts
    n: tuple
    L: tuple
    P: int
    physics: object
    bcs: object
    limiter: bool
    eps: float
    dt: float
    nsteps: int
    ic: Callable[[np.ndarray], np.ndarray]
    params: dict = field(default_factory=dict)


def dam_break_ic(xy, ratio, nc=3):
    """Shallow-water dam break at x = 0.5: h = ratio left, 1 right, at rest
    (examples/shallow_water_dam_break_2d_gpu.mojo:70-80)."""
    h = np.where(xy[..., 0] < 0.5, ratio, 1.0)
    return np.stack([h, np.zeros_like(h), np.zeros_like(h)], axis=-1)


def sod_ic(xy, n, gamma=1.4):
    """Sod tube along x on the unit square, (rho, p) = (1, 1) | (0.125, 0.1),
    tanh-smoothed over 8 cells (examples/euler_sod_2d_gpu.mojo:86-104);
    trait Euler state (rho, rhou, rhov, rhow, E)."""
    s = 0.5 * (np.tanh((xy[..., 0] - 0.5) / (8.0 / n)) + 1.0)
    rho = 1.0 + s * (0.125 - 1.0)
    p = 1.0 + s * (0.1 - 1.0)
    z = np.zeros_like(rho)
    return np.stack([rho, z, z, z, p / (gamma - 1.0)], axis=-1)


def cavity_ic(xy, m, k):
    """TM(m, k) mode of the PEC unit square at t = 0: Ez = sin(m pi x)
    sin(k pi y), B = 0 (examples/maxwell_cavity_2d_gpu.mojo:8-10)."""
    ez = np.sin(m * math.pi * xy[..., 0]) * np.sin(k * math.pi * xy[..., 1])
    z = np.zeros_like(ez)
    return np.stack([z, z, ez, z, z, z], axis=-1)


def serving_stream(count=64, rate=20.0, seed=0, n_min=16, n_max=256, orders=(2, 3, 4), t_range=(0.1, 1.0)):
    """The seeded serving stream: Poisson arrivals at `rate` requests/s;
    each request uniformly one of a shallow-water dam break (height ratio
    U[2,10], slip walls, HLL + limiter), an Euler Sod tube (HLLE + limiter,
    outflow x, walls y) or a Maxwell PEC cavity mode (m, k) in {1..4}^2;
    mesh side n ~ U{n_min..n_max} (2 n^2 triangles on the unit square),
    P ~ U(orders), end time ~ U(t_range) = U[0.1, 1]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        kind = ("dam_break", "sod", "cavity")[int(rng.integers(0, 3))]
        n = int(rng.integers(n_min, n_max + 1))
        P = int(orders[int(rng.integers(0, len(orders)))])
        T = float(rng.uniform(*t_range))
        h = 1.0 / n
        if kind == "dam_break":
            ratio = float(rng.uniform(2.0, 10.0))
            phys = ph.ShallowWater2D(g=9.81, hll=True)
            bcs = BoundaryConditions2D(1, 1, 1, 1)
            nsteps, dt = fixed_steps(T, apriori_dt(0.2, h, math.sqrt(9.81 * ratio), P))
            ic = (lambda xy, r=ratio: dam_break_ic(xy, r))
            prm, lim = dict(ratio=ratio), True
        elif kind == "sod":
            phys = ph.Euler(gamma=1.4, flux_type=ph.FLUX_HLLE)
            bcs = BoundaryConditions2D(2, 2, 1, 1)
            nsteps, dt = fixed_steps(T, apriori_dt(0.15, h, 2.0 * math.sqrt(1.4), P))
            ic = (lambda xy, n=n: sod_ic(xy, n))
            prm, lim = {}, True
        else:
            m, k = int(rng.integers(1, 5)), int(rng.integers(1, 5))
            phys = ph.Maxwell2D(c=1.0)
            bcs = BoundaryConditions2D(1, 1, 1, 1)
            nsteps, dt = fixed_steps(T, apriori_dt(0.2, h, 1.0, P))
            ic = (lambda xy, m=m, k=k: cavity_ic(xy, m, k))
            prm, lim = dict(mode=(m, k)), False
        prm.update(t_final=T)
        out.append(Request(i, kind, t, (n, n), (1.0, 1.0), P, phys, bcs, lim, 0.1, dt, nsteps, ic, prm))
    return out
