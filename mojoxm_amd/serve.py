"""Serving loop: continuous batching of solve requests on one GPU.

Requests arrive over time (problems.serving_stream).  The loop admits an
arrived request as soon as its exact device footprint
(mxm_batch_footprint_*) fits the HBM byte budget next to the requests in
flight, adds it to the running batch (it joins at the next step), advances
every request in flight together in chunks of SSPRK3 steps, and retires
finished ones -- their memory is reclaimed and waiting requests move in.
Requests that use the same stage kernel share each launch, so their work
overlaps on the device.  No counterpart in the reference, which runs its
drivers one problem at a time; a request's result is the same, bit for
bit, alone (Solver) and in the mix.

serve() reports per-request time-to-solution (finish - arrival), p50/p99
and solves/s; one_at_a_time() runs the same requests sequentially, each
with its own Solver, for the baseline."""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .batch import Batch, footprint
from .mesh import LocalMesh2D
from .solver import PREC_FP64, Solver


@dataclass
class ServeReport:
    solves: int
    makespan: float          # first arrival to last completion, s
    latencies: list          # time-to-solution per request, s
    peak_bytes: int          # largest device footprint in flight
    budget: int
    results: dict = field(default_factory=dict)

    @property
    def solves_per_s(self) -> float:
        return self.solves / self.makespan

    def pct(self, q: float) -> float:
        return float(np.percentile(self.latencies, q))


def initial_state(req):
    mesh = LocalMesh2D(*req.n, *req.L, bcs=req.bcs, P=req.P)
    return req.ic(mesh.elem_node_xyz)


def serve(requests, precision=PREC_FP64, budget=64 << 30, chunk=16, keep_results=False,
          states=None) -> ServeReport:
    """Run the stream with continuous batching.  `chunk` SSPRK3 steps run
    between admission/retirement points; `states` may hold precomputed
    initial states by request index."""
    reqs = sorted(requests, key=lambda r: r.arrival)
    fps = {r.index: footprint(r.n, r.L, r.physics, r.P, precision, r.limiter, r.bcs) for r in reqs}
    if max(fps.values()) > budget:
        raise ValueError("a request alone exceeds the HBM budget")
    q0 = states or {r.index: initial_state(r) for r in reqs}
    pending = list(reqs)
    inflight = {}  # batch id -> request
    used = peak = 0
    lat, out = {}, {}
    with Batch(precision) as b:
        t0 = time.perf_counter()
        while pending or inflight:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now and used + fps[pending[0].index] <= budget:
                r = pending.pop(0)
                i = b.add(r.n, r.L, r.physics, r.P, q0[r.index], r.dt, r.nsteps, bcs=r.bcs, limiter=r.limiter,
                          eps=r.eps)
                inflight[i] = r
                used += fps[r.index]
                peak = max(peak, used)
            if not inflight:
                # idle until the next arrival
                time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
                continue
            b.step(chunk)
            done = b.poll()
            now = time.perf_counter() - t0
            for i in done:
                r = inflight.pop(i)
                lat[r.index] = now - r.arrival
                if keep_results:
                    out[r.index] = b.get_state(i)
                b.remove(i)
                used -= fps[r.index]
        makespan = time.perf_counter() - t0 - reqs[0].arrival
    return ServeReport(len(reqs), makespan, [lat[r.index] for r in reqs], peak, budget, out)


def one_at_a_time(requests, precision=PREC_FP64, states=None, keep_results=False, tune=False):
    """Each request alone with its own Solver, in arrival order.  Returns
    (solve seconds per request, results); time-to-solution of a FIFO server
    follows from the arrivals: finish_i = max(arrival_i, finish_{i-1}) + t_i.
    tune: autotune each solver's launch configuration first (untimed)."""
    q0 = states or {r.index: initial_state(r) for r in requests}
    secs, out = {}, {}
    for r in sorted(requests, key=lambda r: r.arrival):
        with Solver(r.n, r.L, r.physics, P=r.P, bcs=r.bcs, precision=precision) as s:
            s.set_state(q0[r.index])
            if r.limiter:
                s.enable_cell_limiter(True, r.eps)
            if tune:
                s.autotune(r.dt, 5)
            s.synchronize()
            t = time.perf_counter()
            s.step_ssprk3(r.dt, r.nsteps)
            s.synchronize()
            secs[r.index] = time.perf_counter() - t
            if keep_results:
                out[r.index] = s.get_state()
    return secs, out


def fifo_latencies(requests, secs):
    """Time-to-solution of a one-at-a-time FIFO server with the given solve
    times."""
    finish, lat = 0.0, []
    for r in sorted(requests, key=lambda r: r.arrival):
        finish = max(r.arrival, finish) + secs[r.index]
        lat.append(finish - r.arrival)
    return lat
