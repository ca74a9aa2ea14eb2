"""Boundary-condition kinds and per-domain configurations
(src/boundary.mojo:24-76).  BC_INTERIOR on a side means periodic."""
from __future__ import annotations

from dataclasses import dataclass

BC_INTERIOR = 0
BC_WALL = 1
BC_OUTFLOW = 2
BC_INFLOW = 3


@dataclass(frozen=True)
class BoundaryConditions:
    """Six-face order [-x, +x, -y, +y, -z, +z]."""
    bc_x_lo: int = BC_INTERIOR
    bc_x_hi: int = BC_INTERIOR
    bc_y_lo: int = BC_INTERIOR
    bc_y_hi: int = BC_INTERIOR
    bc_z_lo: int = BC_INTERIOR
    bc_z_hi: int = BC_INTERIOR

    @staticmethod
    def periodic() -> "BoundaryConditions":
        return BoundaryConditions()

    def all_periodic(self) -> bool:
        return all(v == BC_INTERIOR for v in self.as_list())

    def as_list(self):
        return [self.bc_x_lo, self.bc_x_hi, self.bc_y_lo, self.bc_y_hi, self.bc_z_lo, self.bc_z_hi]


@dataclass(frozen=True)
class BoundaryConditions2D:
    """Four-face order [-x, +x, -y, +y]."""
    bc_x_lo: int = BC_INTERIOR
    bc_x_hi: int = BC_INTERIOR
    bc_y_lo: int = BC_INTERIOR
    bc_y_hi: int = BC_INTERIOR

    @staticmethod
    def periodic() -> "BoundaryConditions2D":
        return BoundaryConditions2D()

    def all_periodic(self) -> bool:
        return all(v == BC_INTERIOR for v in self.as_list())

    def as_list(self):
        return [self.bc_x_lo, self.bc_x_hi, self.bc_y_lo, self.bc_y_hi]
