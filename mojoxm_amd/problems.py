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
