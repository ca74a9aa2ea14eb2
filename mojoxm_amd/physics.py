"""Physics descriptors: the reference's Physics types as parameter records
that select a compiled stage kernel (ids and layouts of include/mxm_dg.h).

Trait physics (usable on tets and triangles):
  Advection  src/advection.mojo:20-119
  Euler      src/euler.mojo:523-792 (flux_type Rusanov/Roe/HLLE/HLLEC)
2D module physics (triangles only):
  Euler2D        src/local_mesh_2d_gpu_euler.mojo      (Rusanov / HLLC)
  ShallowWater2D src/local_mesh_2d_gpu_sw.mojo         (Rusanov / HLL)
  MHDGLM2D       src/local_mesh_2d_gpu_mhd_glm.mojo    (Rusanov, psi damping)
  Maxwell2D      src/local_mesh_2d_gpu_maxwell.mojo    (Rusanov, PEC walls)
  Advection2D    src/local_mesh_2d_gpu_advection.mojo  (upwind)
Other trait physics of the 3D solver (tets only, Rusanov fluxes):
  ShallowWater        src/shallow_water.mojo (2D system embedded in 3D)
  Maxwell             src/maxwell.mojo (PEC walls, J and M sources)
  IdealMHD            src/mhd.mojo (GLM cleaning, damping as a source)
  FiveMomentTwoFluid  src/two_fluid.mojo (electron + ion fluids, fields)
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

(PHYS_ADVECTION, PHYS_EULER, PHYS_EULER2D, PHYS_SW2D, PHYS_MHDGLM2D, PHYS_MAXWELL2D, PHYS_ADVECTION2D, PHYS_SW, PHYS_MAXWELL,
 PHYS_MHD, PHYS_TWOFLUID) = range(11)
FLUX_RUSANOV, FLUX_ROE, FLUX_HLLE, FLUX_HLLEC = range(4)


def _pack(vals) -> np.ndarray:
    out = np.zeros(32)
    v = np.asarray(vals, dtype=np.float64)
    out[: v.size] = v
    return out


@dataclass
class Advection:
    vx: float
    vy: float
    vz: float = 0.0
    inflow_q: float = 0.0
    phys_id = PHYS_ADVECTION
    NUM_COMPONENTS = 1
    dims = (2, 3)

    def params(self) -> np.ndarray:
        return _pack([self.vx, self.vy, self.vz, self.inflow_q])


@dataclass
class Euler:
    gamma: float = 1.4
    min_density: float = 1e-6
    min_pressure: float = 1e-6
    flux_type: int = FLUX_RUSANOV
    entropy_fix: bool = False
    gx: float = 0.0
    gy: float = 0.0
    gz: float = 0.0
    inflow: tuple = (0.0, 0.0, 0.0, 0.0, 0.0)
    phys_id = PHYS_EULER
    NUM_COMPONENTS = 5
    dims = (2, 3)

    def params(self) -> np.ndarray:
        return _pack([self.gamma, self.min_density, self.min_pressure, self.flux_type, float(self.entropy_fix),
                      self.gx, self.gy, self.gz, *self.inflow])


@dataclass
class Euler2D:
    gamma: float = 1.4
    min_density: float = 1e-6
    min_pressure: float = 1e-6
    hllc: bool = False
    gx: float = 0.0
    gy: float = 0.0
    inflow: tuple = (0.0, 0.0, 0.0, 0.0)
    phys_id = PHYS_EULER2D
    NUM_COMPONENTS = 4
    dims = (2,)

    def params(self) -> np.ndarray:
        return _pack([self.gamma, self.min_density, self.min_pressure, float(self.hllc), self.gx, self.gy,
                      *self.inflow])


@dataclass
class ShallowWater2D:
    g: float = 9.81
    min_h: float = 1e-6
    hll: bool = False
    inflow: tuple = (0.0, 0.0, 0.0)
    phys_id = PHYS_SW2D
    NUM_COMPONENTS = 3
    dims = (2,)

    def params(self) -> np.ndarray:
        return _pack([self.g, self.min_h, float(self.hll), *self.inflow])


@dataclass
class MHDGLM2D:
    gamma: float = 5.0 / 3.0
    min_density: float = 1e-6
    min_pressure: float = 1e-6
    c_h: float = 1.0
    alpha_d: float = 0.0
    inflow: tuple = (0.0,) * 7
    phys_id = PHYS_MHDGLM2D
    NUM_COMPONENTS = 7
    dims = (2,)

    def params(self) -> np.ndarray:
        return _pack([self.gamma, self.min_density, self.min_pressure, self.c_h, self.alpha_d, *self.inflow])


@dataclass
class Maxwell2D:
    c: float = 1.0
    J: tuple = (0.0, 0.0, 0.0)
    M: tuple = (0.0, 0.0, 0.0)
    inflow: tuple = (0.0,) * 6
    phys_id = PHYS_MAXWELL2D
    NUM_COMPONENTS = 6
    dims = (2,)

    def params(self) -> np.ndarray:
        return _pack([self.c, *self.J, *self.M, *self.inflow])


@dataclass
class Advection2D:
    vx: float
    vy: float
    inflow_q: float = 0.0
    phys_id = PHYS_ADVECTION2D
    NUM_COMPONENTS = 1
    dims = (2,)

    def params(self) -> np.ndarray:
        return _pack([self.vx, self.vy, self.inflow_q])


@dataclass
class ShallowWater:
    """src/shallow_water.mojo:31-52: state (h, hu, hv), F^z = 0."""
    g: float = 1.0
    h_min: float = 1e-6
    inflow: tuple = (0.0, 0.0, 0.0)
    phys_id = PHYS_SW
    NUM_COMPONENTS = 3
    dims = (3,)

    def params(self) -> np.ndarray:
        return _pack([self.g, self.h_min, *self.inflow])


@dataclass
class Maxwell:
    """src/maxwell.mojo:48-110: state (E, B), uniform sources J and M."""
    c: float = 1.0
    J: tuple = (0.0, 0.0, 0.0)
    M: tuple = (0.0, 0.0, 0.0)
    inflow: tuple = (0.0,) * 6
    phys_id = PHYS_MAXWELL
    NUM_COMPONENTS = 6
    dims = (3,)

    def params(self) -> np.ndarray:
        return _pack([self.c, *self.J, *self.M, *self.inflow])


@dataclass
class IdealMHD:
    """src/mhd.mojo:197-252: state (rho, rho u[3], E, B[3], psi); GLM with
    speed c_h and damping alpha_d (a source term)."""
    gamma: float = 5.0 / 3.0
    min_density: float = 1e-6
    min_pressure: float = 1e-6
    c_h: float = 0.0
    alpha_d: float = 0.0
    inflow: tuple = (0.0,) * 9
    phys_id = PHYS_MHD
    NUM_COMPONENTS = 9
    dims = (3,)

    def params(self) -> np.ndarray:
        return _pack([self.gamma, self.min_density, self.min_pressure, self.c_h, self.alpha_d, *self.inflow])


@dataclass
class FiveMomentTwoFluid:
    """src/two_fluid.mojo:140-210: state (electron rho, m[3], E; ion rho,
    m[3], E; field E[3], B[3], psi)."""
    gamma_e: float = 5.0 / 3.0
    gamma_i: float = 5.0 / 3.0
    q_e: float = -1.0
    m_e: float = 1.0 / 25.0
    q_i: float = 1.0
    m_i: float = 1.0
    eps0: float = 1.0
    c_light: float = 1.0
    c_h: float = 0.0
    alpha_d: float = 0.0
    min_density: float = 1e-6
    min_pressure: float = 1e-6
    inflow: tuple = (0.0,) * 17
    phys_id = PHYS_TWOFLUID
    NUM_COMPONENTS = 17
    dims = (3,)

    def params(self) -> np.ndarray:
        return _pack([self.gamma_e, self.gamma_i, self.q_e, self.m_e, self.q_i, self.m_i, self.eps0, self.c_light,
                      self.c_h, self.alpha_d, self.min_density, self.min_pressure, *self.inflow])
