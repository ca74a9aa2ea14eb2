"""SSPRK3 stage plans (src/ssprk3.mojo:41-78): the Shu-Osher routing of the
three stages over the d_q / d_q1 / d_q2 buffer trio."""
from __future__ import annotations

from dataclasses import dataclass


@dataclass(frozen=True)
class SSPRK3StagePlan:
    q_in: str
    q_a: str
    q_b: str
    q_out: str
    a: float
    b: float
    c: float


def ssprk3_stage_plans(d_q="q", d_q1="q1", d_q2="q2"):
    """q1 = q + dt L(q); q2 = 3/4 q + 1/4 (q1 + dt L(q1)); q = 1/3 q + 2/3 (q2 + dt L(q2))."""
    return [
        SSPRK3StagePlan(d_q, d_q, d_q, d_q1, 1.0, 0.0, 1.0),
        SSPRK3StagePlan(d_q1, d_q, d_q1, d_q2, 0.75, 0.25, 0.25),
        SSPRK3StagePlan(d_q2, d_q, d_q2, d_q, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0),
    ]
