"""mojoxm_amd -- MI355X (gfx950) implementation of mojoxm's DG right-hand
side and SSPRK3 time step: hand-written HIP stage kernels behind the C ABI
of libmxm_hip.so (include/mxm_dg.h), with this package as the host-side
mirror of the reference's Solver / driver interface."""
from . import batch, diagnostics, mxmq, physics, problems, serve, vtu
from ._lib import MxmError
from .diagnostics import DiagComponents, Diagnostics, DiagnosticsWriter
from .boundary import (BC_INFLOW, BC_INTERIOR, BC_OUTFLOW, BC_WALL, BoundaryConditions,
                       BoundaryConditions2D)
from .memory_report import MemoryReport, ThroughputReport
from .mesh import LocalMesh, LocalMesh2D
from .reference import ReferenceElement, ReferenceElement2D, num_tet_nodes, num_tri_nodes
from .solver import PREC_FP32_MFMA, PREC_FP64, Solver
from .ssprk3 import SSPRK3StagePlan, ssprk3_stage_plans
from .time_integrator import (TimeLoopResult, apriori_dt, cfl_dt, fixed_steps, run_ssprk3_loop,
                              run_ssprk3_loop_with_diagnostics)

__all__ = [
    "batch", "diagnostics", "mxmq", "serve", "DiagComponents", "Diagnostics", "DiagnosticsWriter",
    "physics", "problems", "MxmError", "BC_INFLOW", "BC_INTERIOR", "BC_OUTFLOW", "BC_WALL",
    "BoundaryConditions", "BoundaryConditions2D", "MemoryReport", "ThroughputReport", "LocalMesh", "LocalMesh2D",
    "ReferenceElement", "ReferenceElement2D", "num_tet_nodes", "num_tri_nodes", "Solver", "SSPRK3StagePlan",
    "ssprk3_stage_plans", "TimeLoopResult", "apriori_dt", "fixed_steps", "run_ssprk3_loop",
    "run_ssprk3_loop_with_diagnostics", "cfl_dt", "vtu",
]
