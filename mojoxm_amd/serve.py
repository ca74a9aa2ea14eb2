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
    solves: int              # requests that completed
    makespan: float          # first arrival to last completion, s
    latencies: list          # time-to-solution per completed request, s
    peak_bytes: int          # largest device footprint in flight (bound incl. a compaction's second arena)
    budget: int
    results: dict = field(default_factory=dict)
    failed: list = field(default_factory=list)  # request indices evicted: a stored value went non-finite

    @property
    def solves_per_s(self) -> float:
        return self.solves / self.makespan

    def pct(self, q: float) -> float:
        return float(np.percentile(self.latencies, q))


def initial_state(req):
    mesh = LocalMesh2D(*req.n, *req.L, bcs=req.bcs, P=req.P)
    return req.ic(mesh.elem_node_xyz)


def _group(req):
    """Requests with the same key may share a kernel group (coarser than the
    library's key, so the compaction reserve below is an upper bound)."""
    return (len(req.n), req.P, req.physics.phys_id, bool(req.limiter))


def serve(requests, precision=PREC_FP64, budget=64 << 30, chunk=16, keep_results=False,
          states=None, headroom=0.5, progress=None) -> ServeReport:
    """Run the stream with continuous batching.  `chunk` SSPRK3 steps run
    between admission/retirement points; `states` may hold precomputed
    initial states by request index.

    The device stays busy while the host works: each round enqueues the
    next chunk first, then admits arrivals (their meshes and tables are
    built on the host meanwhile) and waits only for the previous chunk
    (Batch.poll(lagged=True)) before retiring what finished.  A request
    whose state goes non-finite is switched off on the device at that step,
    reported failed and evicted; its memory returns to the budget.

    Admission: the device bytes held now (Batch.memory(), plus the
    footprints of requests admitted since the last step), plus the new
    request's footprint, plus the footprint of its whole kernel group
    (a compaction holds the old and the new arena of that group at once)
    must fit the budget.  An arena that has had to grow is rebuilt with
    `headroom` (a fraction) of spare tiles and slots, so that later
    arrivals are placed without another rebuild; the spare memory counts
    against the budget like the rest.  progress(done, n), if given, is
    called as requests complete or fail."""
    reqs = sorted(requests, key=lambda r: r.arrival)
    fps = {r.index: footprint(r.n, r.L, r.physics, r.P, precision, r.limiter, r.bcs) for r in reqs}
    if max(2 * f for f in fps.values()) > budget:
        raise ValueError("a request alone (with its compaction copy) exceeds the HBM budget")
    q0 = states or {r.index: initial_state(r) for r in reqs}
    pending = list(reqs)
    inflight = {}  # batch id -> request
    group_fp = {}  # kernel group -> footprint of its requests in flight
    peak = 0
    lat, out, failed = {}, {}, []
    with Batch(precision) as b:
        b.headroom(headroom)
        t0 = time.perf_counter()
        unplaced = 0  # footprints added since the last step
        stepped = False
        while pending or inflight:
            if inflight:
                b.step(chunk)
                unplaced = 0
                stepped = True
            # admit what has arrived and fits
            now = time.perf_counter() - t0
            held = b.memory() + unplaced
            while pending and pending[0].arrival <= now:
                r = pending[0]
                f, key = fps[r.index], _group(r)
                bound = held + f + group_fp.get(key, 0) + f
                if bound > budget:
                    break
                pending.pop(0)
                i = b.add(r.n, r.L, r.physics, r.P, q0[r.index], r.dt, r.nsteps, bcs=r.bcs, limiter=r.limiter,
                          eps=r.eps)
                inflight[i] = r
                group_fp[key] = group_fp.get(key, 0) + f
                held += f
                unplaced += f
                peak = max(peak, bound)
            if not inflight:
                # idle until the next arrival
                time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
                continue
            if not stepped:
                continue
            done = b.poll(lagged=True)
            now = time.perf_counter() - t0
            peak = max(peak, b.memory())
            for i in done:
                r = inflight.pop(i)
                if b.failed(i):
                    failed.append(r.index)
                else:
                    lat[r.index] = now - r.arrival
                    if keep_results:
                        out[r.index] = b.get_state(i)
                b.remove(i)
                group_fp[_group(r)] -= fps[r.index]
                if progress:
                    progress(len(lat) + len(failed), len(reqs))
        makespan = time.perf_counter() - t0 - reqs[0].arrival
    ok = [r for r in reqs if r.index in lat]
    return ServeReport(len(ok), makespan, [lat[r.index] for r in ok], peak, budget, out, failed)


def one_at_a_time(requests, precision=PREC_FP64, states=None, keep_results=False, tune=False, setup=None,
                  failed=None, check_every=512, progress=None):
    """Each request alone with its own Solver, in arrival order.  Returns
    (solve seconds per request, results): the SSPRK3 stepping only.  Pass a
    dict as `setup` to also get each request's set-up seconds (Solver
    creation -- mesh, tables, device buffers -- the state upload and the
    autotuning), which a real one-at-a-time server pays too.  A FIFO
    server's time-to-solution follows from the arrivals:
    finish_i = max(arrival_i, finish_{i-1}) + t_i.
    tune: autotune each solver's launch configuration first.
    Like the served requests, a request stops when its state is no longer
    finite: every `check_every` steps the device speed reduction (NaN or
    infinite for a non-finite state) is read, and the indices of requests
    stopped that way go to the list `failed`.  progress(k, n), if given,
    is called after each request."""
    q0 = states or {r.index: initial_state(r) for r in requests}
    secs, out = {}, {}
    for r in sorted(requests, key=lambda r: r.arrival):
        t_setup = time.perf_counter()
        with Solver(r.n, r.L, r.physics, P=r.P, bcs=r.bcs, precision=precision) as s:
            s.set_state(q0[r.index])
            if r.limiter:
                s.enable_cell_limiter(True, r.eps)
            if tune:
                s.autotune(r.dt, 5)
                s.set_state(q0[r.index])
            chunk = min(check_every, r.nsteps) if check_every else r.nsteps
            s.prepare_steps(r.dt, chunk)
            if r.nsteps % chunk:
                s.prepare_steps(r.dt, r.nsteps % chunk)
            s.synchronize()
            t = time.perf_counter()
            if setup is not None:
                setup[r.index] = t - t_setup
            done = 0
            while done < r.nsteps:
                k = min(chunk, r.nsteps - done)
                s.step_ssprk3(r.dt, k)
                done += k
                if check_every and not np.isfinite(s.max_wave_speed()):
                    if failed is not None:
                        failed.append(r.index)
                    break
            s.synchronize()
            secs[r.index] = time.perf_counter() - t
            if keep_results:
                out[r.index] = s.get_state()
        if progress:
            progress(len(secs), len(requests))
    return secs, out


def fifo_latencies(requests, secs):
    """Time-to-solution of a one-at-a-time FIFO server with the given solve
    times."""
    finish, lat = 0.0, []
    for r in sorted(requests, key=lambda r: r.arrival):
        finish = max(r.arrival, finish) + secs[r.index]
        lat.append(finish - r.arrival)
    return lat
