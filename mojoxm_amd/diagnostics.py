"""Domain diagnostics (src/diagnostics.mojo): mass-weighted conserved
integrals, squares, max |q|, kinetic energy and enstrophy of a
device-resident state, one device call per record (mxm_dg_diagnostics /
mxm_batch_diagnostics), and the reference's CSV writer on top.

The reference's DiagnosticsWriter.record (src/diagnostics.mojo:161-223)
downloads each configured component and sums raw nodal values on the host,
scaled by domain_volume / total_dof -- an unweighted nodal mean (:170).
Here the integrals are mass-weighted (V_e sum_i w_i f(q_i) with the node
weights) and summed on the device in compensated FP64; the CSV layout is
the reference's: time, then the linear, squared and max-abs columns in the
order they were configured, then any derived columns (kinetic_energy,
enstrophy) that were asked for."""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Diagnostics:
    integral: np.ndarray  # int q_c dV, mass-weighted
    squared: np.ndarray   # int q_c^2 dV, mass-weighted
    maxabs: np.ndarray    # max |q_c| over the nodes
    kinetic_energy: float  # int 1/2 rho |u|^2 dV (NaN without a velocity)
    enstrophy: float       # int 1/2 rho |curl u|^2 dV (NaN without a velocity)

    @classmethod
    def from_flat(cls, v, nc: int) -> "Diagnostics":
        v = np.asarray(v, dtype=np.float64)
        return cls(v[:nc].copy(), v[nc:2 * nc].copy(), v[2 * nc:3 * nc].copy(), float(v[3 * nc]),
                   float(v[3 * nc + 1]))

    def flat(self) -> np.ndarray:
        return np.concatenate([self.integral, self.squared, self.maxabs, [self.kinetic_energy, self.enstrophy]])


@dataclass
class DiagComponents:
    """Builder of the labelled component lists (src/diagnostics.mojo:66-92)."""
    linear_list: list = field(default_factory=list)
    squared_list: list = field(default_factory=list)
    maxabs_list: list = field(default_factory=list)
    derived_list: list = field(default_factory=list)

    def linear(self, name: str, c: int) -> "DiagComponents":
        self.linear_list.append((name, c))
        return self

    def squared(self, name: str, c: int) -> "DiagComponents":
        self.squared_list.append((name, c))
        return self

    def maxabs(self, name: str, c: int) -> "DiagComponents":
        self.maxabs_list.append((name, c))
        return self

    def kinetic_energy(self, name: str = "kinetic_energy") -> "DiagComponents":
        self.derived_list.append((name, "kinetic_energy"))
        return self

    def enstrophy(self, name: str = "enstrophy") -> "DiagComponents":
        self.derived_list.append((name, "enstrophy"))
        return self


class DiagnosticsWriter:
    """One CSV row per record(t) (src/diagnostics.mojo:95-223).  `source` is
    a Solver, or (Batch, instance id)."""

    def __init__(self, source, csv_path: str, components: DiagComponents):
        self.source = source
        self.csv_path = csv_path
        self.components = components
        self.last_row: list = []
        d = os.path.dirname(csv_path)
        if d:
            os.makedirs(d, exist_ok=True)
        cols = ["time"] + [n for n, _ in components.linear_list] + [n for n, _ in components.squared_list] + \
            [n for n, _ in components.maxabs_list] + [n for n, _ in components.derived_list]
        with open(csv_path, "w") as f:
            f.write(",".join(cols) + "\n")

    def _diagnostics(self) -> Diagnostics:
        if isinstance(self.source, tuple):
            batch, ident = self.source
            return batch.diagnostics(ident)
        return self.source.diagnostics()

    def record(self, t: float) -> list:
        d = self._diagnostics()
        c = self.components
        row = [float(d.integral[k]) for _, k in c.linear_list] + [float(d.squared[k]) for _, k in c.squared_list] + \
            [float(d.maxabs[k]) for _, k in c.maxabs_list] + [float(getattr(d, k)) for _, k in c.derived_list]
        self.last_row = row
        with open(self.csv_path, "a") as f:
            f.write(",".join(repr(v) for v in [float(t)] + row) + "\n")
        return row
