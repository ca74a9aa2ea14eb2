"""Benchmark of the DG right-hand side + SSPRK3 step on one MI355X.

Workload: 2D isentropic vortex (Shu form, beta=5, gamma=1.4, background
velocity (1,1)) on periodic [0,10]^2, 128x128 squares split into 32768
triangles, P=3 (NP=10), trait Euler NC=5, HLLEC flux, run to T=2
(4054 SSPRK3 steps at CFL 0.15) -- the configuration the metric is quoted
on.  One "step" is one SSPRK3 step (three stage kernels) over the whole
mesh, state resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  `value` is the DOF-stage-updates/s of the FP32-MFMA
form; the plain FP64 form is timed beside it (config.fp64, the speedup and
roofline_fp64).  Both forms state their relative L2 and Linf error against
the CPU FP64 oracle (bounded sample of the same run) and their L2 error
against the analytic vortex at T=2; the FP32 form also against the FP64
form at T=2.  roofline: algorithmic bytes / HIP-event kernel time for the
stage kernel.  cpu_baseline: the oracle's throughput on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
METRIC = "DOF-stage-updates/s FP32-MFMA vs FP64 @ stated L2; batched solves/s + p99"


# Summaries of the rocprofv3 counter passes over this workload's stage
# kernels (one per arithmetic form); their "traffic_bytes_per_launch:" line
# gives roofline.traffic.
TRAFFIC_PROFILES = {
    "fp64": os.path.join(ROOT, "profiles", "dg_stage_kernel_euler_vortex_2d_p3.md"),
    "fp32_mfma": os.path.join(ROOT, "profiles", "dg_stage_kernel_fp32_mfma_euler_vortex_2d_p3.md"),
}
# the same kernels on the 512^2 vortex (both forms in one summary, one
# "traffic_bytes_per_launch_<form>:" line each)
LARGE_PROFILE = os.path.join(ROOT, "profiles", "dg_stage_kernel_euler_vortex_2d_n512_p3.md")


# occupancy target (waves per SIMD) of each launch configuration; 1 = the
# compiler's own register budget
CFG_WAVES = {"fp64": ["compiler", "5", "4"], "fp32_mfma": ["compiler", "5", "6", "7"]}


def tile_report(s):
    """The tiling the autotuner kept: elements per tile, threads per block."""
    wide, elems, threads = s.tile
    return {"elements": elems, "threads": threads, "wide": wide}


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def measured_traffic(form, path=None):
    """HBM bytes per stage launch from a committed counter summary, or None:
    the C2 summary of `form`, or the line of `form` in the summary at path."""
    key = "traffic_bytes_per_launch" if path is None else f"traffic_bytes_per_launch_{form}"
    try:
        with open(path or TRAFFIC_PROFILES[form]) as f:
            m = re.search(rf"^{key}:\s*([0-9.eE+]+)", f.read(), re.M)
    except OSError:
        return None
    return float(m.group(1)) if m else None


def copy_calibration() -> int:
    """Known-size copy for counter calibration: 1 GiB read + 1 GiB written per
    launch, four launches.  Its FETCH_SIZE / WRITE_SIZE readings give the
    factor that turns the counters' KiB into HBM bytes."""
    import torch

    x = torch.ones(2 ** 27, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(4):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read_per_launch": x.numel() * 8, "copy_bytes_written_per_launch": y.numel() * 8,
                      "launches": 4}))
    return 0


def batched_leg(M, args):
    """Seeded ensemble of 2D GLM-MHD Alfven waves (P=3, NC=7, n log-uniform
    in [16, 96], T=1), all instances in one batch, against the same
    instances solved one at a time (bounded, evenly spaced sample); both
    arithmetic forms, the FP32-MFMA one first (the headline)."""
    ens = M.problems.alfven_ensemble(count=args.ensemble)
    q0 = {}
    for m in ens:
        mesh = M.LocalMesh2D(*m.n, *m.L, P=m.P)
        q0[m.index] = m.ic(mesh.elem_node_xyz).astype(np.float32).astype(np.float64)
    work = sum(3.0 * 2 * m.n[0] * m.n[1] * 10 * 7 * m.nsteps for m in ens)

    def run_batch(prec):
        finish = {}
        with M.batch.Batch(prec) as b:
            ids = {b.add(m.n, m.L, m.physics, m.P, q0[m.index], m.dt, m.nsteps): m for m in ens}
            b.step(0)  # every state resident in HBM before the clock starts
            b.synchronize()
            t0 = time.perf_counter()
            # the next chunk is queued before the host waits for the one
            # before it (lagged poll), so the device does not idle
            while len(finish) < len(ens):
                b.step(50)
                done = b.poll(lagged=True)
                now = time.perf_counter() - t0
                for i in done:
                    finish[i] = now
            wall = time.perf_counter() - t0
            states = {i: b.get_state(i) for i in ids}
            mem = b.memory()
        return ids, finish, wall, states, mem

    ids, finish, wall, states, mem = run_batch(M.PREC_FP32_MFMA)
    exact_err = []
    for i, m in ids.items():
        mesh = M.LocalMesh2D(*m.n, *m.L, P=m.P)
        exact_err.append(rel_l2(states[i], m.exact(mesh.elem_node_xyz, m.dt * m.nsteps)))
    _, finish64, wall64, _, _ = run_batch(M.PREC_FP64)
    # one at a time, FP32-MFMA and FP64, on an evenly spaced sample
    step = max(1, len(ens) // args.solo_sample)
    sample = ens[::step][: args.solo_sample]
    solo = {M.PREC_FP32_MFMA: 0.0, M.PREC_FP64: 0.0}
    err64 = []
    for m in sample:
        outs = {}
        for prec in (M.PREC_FP32_MFMA, M.PREC_FP64):
            with M.Solver(m.n, m.L, m.physics, P=m.P, precision=prec) as s:
                s.set_state(q0[m.index])
                s.autotune(m.dt, 5)
                s.prepare_steps(m.dt, m.nsteps)
                # a warm device, as in solves run back to back (the batch
                # runs without a break): about 30 ms of untimed steps, then
                # the initial state again
                s.time_stage_kernel(m.dt, 10)
                s.set_state(q0[m.index])
                s.synchronize()
                t = time.perf_counter()
                s.step_ssprk3(m.dt, m.nsteps)
                s.synchronize()
                solo[prec] += time.perf_counter() - t
                outs[prec] = s.get_state()
        err64.append(rel_l2(outs[M.PREC_FP32_MFMA], outs[M.PREC_FP64]))
    solo_rate = len(sample) / solo[M.PREC_FP32_MFMA]
    solo_rate64 = len(sample) / solo[M.PREC_FP64]
    lat, lat64 = sorted(finish.values()), sorted(finish64.values())
    return {"workload": f"{len(ens)} GLM-MHD Alfven-wave instances (seed 0), P=3, NC=7, n in [16,96], T=1",
            "precision": "fp32_mfma", "solves": len(ens), "wall_s": wall, "solves_per_s": len(ens) / wall,
            "dof_stage_updates_per_s": work / wall, "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)), "device_bytes": mem,
            "one_at_a_time_solves_per_s": solo_rate,
            "one_at_a_time_sample": f"{len(sample)} evenly spaced instances, each alone with its own Solver "
                                    "(autotuned, graphs prepared and the device warmed up before the timed solve)",
            "speedup_vs_one_at_a_time": (len(ens) / wall) / solo_rate,
            "fp64": {"wall_s": wall64, "solves_per_s": len(ens) / wall64, "p50_s": float(np.percentile(lat64, 50)),
                     "p99_s": float(np.percentile(lat64, 99)), "one_at_a_time_solves_per_s": solo_rate64,
                     "speedup_vs_one_at_a_time": (len(ens) / wall64) / solo_rate64},
            "l2_vs_linear_wave_mean": float(np.mean(exact_err)), "l2_vs_linear_wave_max": float(np.max(exact_err)),
            "l2_fp32_vs_fp64_sample_max": float(np.max(err64))}


def advection_leg(M, args):
    """C1: 3D advection of a Gaussian pulse, 16^3 cubes x 6 tets, P=2, NC=1,
    upwind, CFL 0.3, 462 SSPRK3 steps to T=1, both forms on the GPU (full
    run), and the oracle's full run on the host on one core and on all
    cores, with each result's L2 against the oracle and the initial pulse
    (the exact solution at T=1)."""
    pb = M.problems.advection_gaussian_3d()
    mesh = M.LocalMesh(*pb.n, *pb.L, P=pb.P)
    q0 = pb.ic(mesh.elem_node_xyz).astype(np.float32).astype(np.float64)
    res, finals = {}, {}
    for name, prec, word in (("fp64", M.PREC_FP64, 8), ("fp32_mfma", M.PREC_FP32_MFMA, 4)):
        with M.Solver(pb.n, pb.L, pb.physics, P=pb.P, precision=prec) as s:
            U = s.dof_count()
            s.set_state(q0)
            cfg, _ = s.autotune(pb.dt, 20)
            tile = tile_report(s)
            s.set_state(q0)
            s.prepare_steps(pb.dt, pb.nsteps)
            s.synchronize()
            t0 = time.perf_counter()
            s.step_ssprk3(pb.dt, pb.nsteps)
            s.synchronize()
            wall = time.perf_counter() - t0
            finals[name] = s.get_state()
            s.set_state(q0)
            avg_us = s.time_stage_kernel(pb.dt, 50)
        alg = 8.0 / 3.0 * U * word
        res[name] = {"value": 3.0 * U * pb.nsteps / wall, "wall_s": wall, "kernel_avg_us": avg_us,
                     "achieved_gbs": alg / (avg_us * 1e-6) / 1e9, "frac": alg / (avg_us * 1e-6) / 1e9 / HBM_PEAK_GBS,
                     "launch_config": {"chosen": cfg, "waves_per_simd": CFG_WAVES[name][cfg],
                                       "tile": tile},
                     "l2_vs_initial_T1": rel_l2(finals[name], q0)}
    out = {"workload": "advection_gaussian_3d: 16^3 x 6 tets, P=2, NC=1, upwind, CFL 0.3, 462 steps to T=1",
           "dof_per_stage": U, "dt": pb.dt, "steps": pb.nsteps, "fp64": res["fp64"], "fp32_mfma": res["fp32_mfma"]}
    if not args.no_cpu:
        from oracle import oracle as O

        om = O.Mesh(3, pb.P, pb.n, pb.L)
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        cpu = {}
        for label, nth in (("1_core", 1), ("all_cores", threads)):
            O.set_threads(nth)
            c0 = time.perf_counter()
            q_cpu = om.step(pb.physics.phys_id, pb.physics.params(), q0, pb.dt, pb.nsteps)
            c_wall = time.perf_counter() - c0
            cpu[label] = {"value": 3.0 * U * pb.nsteps / c_wall, "unit": "DOF-stage-updates/s", "cores": nth,
                          "wall_s": c_wall}
        O.set_threads(threads)
        out["cpu_oracle_full_run"] = cpu
        out["l2_fp64_vs_oracle_T1"] = rel_l2(finals["fp64"], q_cpu)
        out["l2_fp32_mfma_vs_oracle_T1"] = rel_l2(finals["fp32_mfma"], q_cpu)
    return out


def profile_batch(M, args) -> int:
    """The batched ensemble for a kernel tracer: the same 512 instances in
    one batch per arithmetic form, every instance `args.profile_batch`
    steps, the kernels launched directly on the streams (no graph replay).
    Prints one JSON line with the steps and the wall time per form."""
    ens = M.problems.alfven_ensemble(count=args.ensemble)
    out = {"workload": f"{len(ens)} GLM-MHD Alfven-wave instances (seed 0), P=3, NC=7, n in [16,96]",
           "steps": args.profile_batch, "launch": "direct stream launches (no graphs)"}
    states = {}
    for m in ens:
        mesh = M.LocalMesh2D(*m.n, *m.L, P=m.P)
        states[m.index] = m.ic(mesh.elem_node_xyz).astype(np.float32).astype(np.float64)
    for name, prec in (("fp32_mfma", M.PREC_FP32_MFMA), ("fp64", M.PREC_FP64)):
        with M.batch.Batch(prec) as b:
            b.graphs(False)
            for m in ens:
                b.add(m.n, m.L, m.physics, m.P, states[m.index], m.dt, args.profile_batch)
            b.step(0)
            b.synchronize()
            t0 = time.perf_counter()
            b.step(args.profile_batch)
            b.synchronize()
            out[name] = {"wall_s": time.perf_counter() - t0, "device_bytes": b.memory()}
    print(json.dumps(out))
    return 0


def large_vortex_leg(M, args, nsteps=None):
    """The C2 vortex problem at 512^2 squares (524,288 triangles, P=3, U =
    26.2 M DOF per stage): the same kernels on a state trio larger than the
    256 MB Infinity Cache in both forms (629 MB FP64, 315 MB FP32), so the
    roofline fraction is against HBM.  Each form is autotuned, then its
    stage kernel timed with HIP events; with nsteps, that many steps are
    also run back to back (the profiling mode)."""
    pb = M.problems.euler_vortex_2d(N=512)
    mesh = M.LocalMesh2D(*pb.n, *pb.L, P=pb.P)
    q0 = pb.ic(mesh.elem_node_xyz).astype(np.float32).astype(np.float64)
    out = {"workload": "euler_vortex_2d: 512x512x2 tris, P=3, NC=5, HLLEC, CFL 0.15"}
    for name, prec, word in (("fp64", M.PREC_FP64, 8), ("fp32_mfma", M.PREC_FP32_MFMA, 4)):
        with M.Solver(pb.n, pb.L, pb.physics, P=pb.P, precision=prec) as s:
            U = s.dof_count()
            s.set_state(q0)
            cfg, cfg_us = s.autotune(pb.dt, 10)
            tile = tile_report(s)
            s.set_state(q0)
            if nsteps:
                s.step_ssprk3(pb.dt, nsteps)
                s.synchronize()
            avg_us = s.time_stage_kernel(pb.dt, 30)
        alg = 8.0 / 3.0 * U * word
        gbs = alg / (avg_us * 1e-6) / 1e9
        out[name] = {"kernel_avg_us": avg_us, "algorithmic_bytes_per_launch": alg, "achieved_gbs": gbs,
                     "frac": gbs / HBM_PEAK_GBS, "traffic": measured_traffic(name, LARGE_PROFILE),
                     "launch_config": {"chosen": cfg, "waves_per_simd": CFG_WAVES[name][cfg],
                                       "us_per_config": dict(zip(CFG_WAVES[name], cfg_us)),
                                       "tile": tile}}
    out["dof_per_stage"] = U
    out["speedup_fp32_mfma_vs_fp64"] = out["fp64"]["kernel_avg_us"] / out["fp32_mfma"]["kernel_avg_us"]
    out["traffic_from"] = os.path.relpath(LARGE_PROFILE, ROOT)
    return out


def profile_limited(M, args) -> int:
    """A limited run for the kernel tracer: the serving stream's Euler Sod
    tube (HLLE, outflow in x, walls in y, Barth-Jespersen limiter with
    Venkatakrishnan eps 0.1) on 192^2 x 2 triangles at P = 3, each form
    autotuned and then stepped args.profile_limited steps.  Prints one
    JSON line with the time per step."""
    from mojoxm_amd.boundary import BoundaryConditions2D
    pr = M.problems
    n, P = 192, 3
    phys = M.physics.Euler(gamma=1.4, flux_type=M.physics.FLUX_HLLE)
    bcs = BoundaryConditions2D(2, 2, 1, 1)
    dt = pr.apriori_dt(0.15, 1.0 / n, 2.0 * np.sqrt(1.4), P)
    mesh = M.LocalMesh2D(n, n, 1.0, 1.0, bcs=bcs, P=P)
    q0 = pr.sod_ic(mesh.elem_node_xyz, n).astype(np.float32).astype(np.float64)
    out = {"workload": f"euler_sod_2d: {n}x{n}x2 tris, P={P}, NC=5, HLLE, BJ limiter (eps 0.1)",
           "steps": args.profile_limited}
    for name, prec in (("fp64", M.PREC_FP64), ("fp32_mfma", M.PREC_FP32_MFMA)):
        with M.Solver((n, n), (1.0, 1.0), phys, P=P, bcs=bcs, precision=prec) as s:
            s.enable_cell_limiter(True, 0.1)
            s.set_state(q0)
            cfg, _ = s.autotune(dt, 10)
            tile = tile_report(s)
            s.set_state(q0)
            s.prepare_steps(dt, args.profile_limited)
            s.synchronize()
            t0 = time.perf_counter()
            s.step_ssprk3(dt, args.profile_limited)
            s.synchronize()
            out[name] = {"ms_per_step": (time.perf_counter() - t0) / args.profile_limited * 1e3,
                         "launch_config": CFG_WAVES[name][cfg], "tile": tile, "dof_per_stage": s.dof_count(),
                         "finite": bool(np.isfinite(s.get_state()).all())}
    print(json.dumps(out))
    return 0


def taylor_green_leg(M, args):
    """C3: compressible Taylor-Green vortex, 32^3 cubes x 6 tets, P=3,
    Rusanov, 200 SSPRK3 steps, FP64 against FP32-MFMA from the same
    FP32-representable initial state; KE and enstrophy tracked every 20
    steps with the device's compensated FP64 diagnostics."""
    pb = M.problems.taylor_green_3d()
    mesh = M.LocalMesh(*pb.n, *pb.L, P=pb.P)
    q0 = pb.ic(mesh.elem_node_xyz).astype(np.float32).astype(np.float64)
    chunk = 20
    res, finals = {}, {}
    for name, prec, word in (("fp64", M.PREC_FP64, 8), ("fp32_mfma", M.PREC_FP32_MFMA, 4)):
        with M.Solver(pb.n, pb.L, pb.physics, P=pb.P, precision=prec) as s:
            U = s.dof_count()
            s.set_state(q0)
            cfg, cfg_us = s.autotune(pb.dt, 10)
            tile = tile_report(s)
            s.set_state(q0)
            s.prepare_steps(pb.dt, pb.nsteps)
            s.synchronize()
            t0 = time.perf_counter()
            s.step_ssprk3(pb.dt, pb.nsteps)
            s.synchronize()
            wall = time.perf_counter() - t0
            # the same run again, stopping every `chunk` steps for diagnostics
            s.set_state(q0)
            d = s.diagnostics()
            hist = [(0, d.kinetic_energy, d.enstrophy)]
            for k in range(chunk, pb.nsteps + 1, chunk):
                s.step_ssprk3(pb.dt, chunk)
                d = s.diagnostics()
                hist.append((k, d.kinetic_energy, d.enstrophy))
            finals[name] = s.get_state()
            s.set_state(q0)
            avg_us = s.time_stage_kernel(pb.dt, 20)
        alg = 8.0 / 3.0 * U * word
        res[name] = {"value": 3.0 * U * pb.nsteps / wall, "ms_per_step": wall / pb.nsteps * 1e3,
                     "kernel_avg_us": avg_us, "achieved_gbs": alg / (avg_us * 1e-6) / 1e9,
                     "frac": alg / (avg_us * 1e-6) / 1e9 / HBM_PEAK_GBS,
                     "launch_config": {"chosen": cfg, "waves_per_simd": CFG_WAVES[name][cfg],
                                       "tile": tile},
                     "ke_enstrophy_every_20_steps": [[k, ke, en] for k, ke, en in hist]}
    h64 = res["fp64"]["ke_enstrophy_every_20_steps"]
    h32 = res["fp32_mfma"]["ke_enstrophy_every_20_steps"]
    return {"workload": "taylor_green_3d: 32^3 x 6 tets, P=3, NC=5, Rusanov, Ma 0.3, 200 steps",
            "dof_per_stage": U, "dt": pb.dt, "fp64": res["fp64"], "fp32_mfma": res["fp32_mfma"],
            "speedup_fp32_mfma_vs_fp64": res["fp32_mfma"]["value"] / res["fp64"]["value"],
            "l2_fp32_vs_fp64_final": rel_l2(finals["fp32_mfma"], finals["fp64"]),
            "ke_rel_diff_final": abs(h32[-1][1] / h64[-1][1] - 1.0),
            "enstrophy_rel_diff_final": abs(h32[-1][2] / h64[-1][2] - 1.0),
            "mfma_and_traffic": "profiles/dg_stage_kernel_tgv_3d_p3.md"}


def serving_leg(M, args):
    """Seeded Poisson stream (dam break / Sod / PEC cavity, n in [16,256],
    P in {2,3,4}, T in [0.1,1]) under continuous batching (FP32-MFMA) with
    an HBM budget, against one-at-a-time FIFO solving of the same stream."""
    reqs = M.problems.serving_stream(count=args.requests, rate=args.rate)
    states = {r.index: M.serve.initial_state(r).astype(np.float32).astype(np.float64) for r in reqs}
    budget = 16 << 30
    rep = M.serve.serve(reqs, M.PREC_FP32_MFMA, budget=budget, chunk=16, states=states)
    out = {"workload": f"{len(reqs)} requests, Poisson {args.rate}/s (seed 0): SW dam break, Euler Sod, "
                       "Maxwell PEC cavity; n in [16,256], P in {2,3,4}, T in [0.1,1]",
           "precision": "fp32_mfma", "hbm_budget_bytes": budget, "peak_bytes": rep.peak_bytes,
           "solves_per_s": rep.solves_per_s, "p50_s": rep.pct(50), "p99_s": rep.pct(99),
           "failed": len(rep.failed)}
    # FIFO one-at-a-time servers of the same stream: FP64 (the stated
    # baseline) and the same FP32-MFMA form, each request's time the
    # stepping alone ("solve") and with its Solver set-up ("with_setup":
    # mesh, tables, buffers, upload, autotuning), which the served
    # makespan also contains
    arrivals = sorted(reqs, key=lambda r: r.arrival)
    out["failed_requests"] = [{"index": r.index, "kind": r.kind, "n": r.n[0], "P": r.P, "steps": r.nsteps}
                              for r in reqs if r.index in rep.failed]
    for name, prec in (("fp64", M.PREC_FP64), ("fp32_mfma", M.PREC_FP32_MFMA)):
        setup, failed = {}, []
        secs, _ = M.serve.one_at_a_time(reqs, prec, states=states, tune=True, setup=setup, failed=failed)
        both = {r.index: secs[r.index] + setup[r.index] for r in reqs}
        ok = [k for k, r in enumerate(arrivals) if r.index not in failed]
        leg = {"failed": len(failed)}
        for label, t in (("solve", secs), ("with_setup", both)):
            fifo = M.serve.fifo_latencies(reqs, t)
            makespan = max(r.arrival + l for r, l in zip(arrivals, fifo)) - arrivals[0].arrival
            lat = [fifo[k] for k in ok]
            leg[label] = {"solves_per_s": len(ok) / makespan, "p50_s": float(np.percentile(lat, 50)),
                          "p99_s": float(np.percentile(lat, 99))}
        out[f"one_at_a_time_{name}"] = leg
        out[f"speedup_vs_one_at_a_time_{name}"] = rep.solves_per_s / leg["solve"]["solves_per_s"]
    return out


def serve_study(M, args) -> int:
    """The serving stream alone, at each arrival rate of args.rates: the
    same seeded requests (args.requests of them) re-timed with each rate's
    Poisson arrivals, served with continuous batching (FP32-MFMA), against
    FIFO one-at-a-time FP32-MFMA solves, whose solve times do not depend on
    the arrivals and are measured once.  Prints one JSON line."""
    base = M.problems.serving_stream(count=args.requests, rate=1.0)
    states = {r.index: M.serve.initial_state(r).astype(np.float32).astype(np.float64) for r in base}
    setup, failed = {}, []
    def progress(k, n):
        print(json.dumps({"one_at_a_time_done": k, "of": n}), flush=True)

    secs, _ = M.serve.one_at_a_time(base, M.PREC_FP32_MFMA, states=states, tune=True, setup=setup, failed=failed,
                                    progress=progress)
    out = {"workload": f"{len(base)} requests (seed 0): SW dam break, Euler Sod, Maxwell PEC cavity; "
                       "n in [16,256], P in {2,3,4}, T in [0.1,1]",
           "precision": "fp32_mfma", "one_at_a_time_failed": sorted(failed),
           "one_at_a_time_solve_s_total": sum(secs.values()), "one_at_a_time_setup_s_total": sum(setup.values()),
           "rates": {}}
    import dataclasses
    for rate in args.rates:
        # the same requests, arrivals rescaled to this rate (Poisson at 1/s scaled)
        reqs = [dataclasses.replace(r, arrival=r.arrival / rate) for r in base]
        def served(k, n, rate=rate):
            print(json.dumps({"rate": rate, "served_done": k, "of": n}), flush=True)

        rep = M.serve.serve(reqs, M.PREC_FP32_MFMA, budget=64 << 30, chunk=16, states=states, progress=served)
        arrivals = sorted(reqs, key=lambda r: r.arrival)
        ok = [k for k, r in enumerate(arrivals) if r.index not in failed]
        leg = {"served": {"solves_per_s": rep.solves_per_s, "solves": rep.solves, "failed": sorted(rep.failed),
                          "p50_s": rep.pct(50), "p99_s": rep.pct(99), "peak_bytes": rep.peak_bytes,
                          "makespan_s": rep.makespan}}
        for label, t in (("solve", secs), ("with_setup", {k: secs[k] + setup[k] for k in secs})):
            fifo = M.serve.fifo_latencies(reqs, t)
            makespan = max(r.arrival + l for r, l in zip(arrivals, fifo)) - arrivals[0].arrival
            lat = [fifo[k] for k in ok]
            leg[f"one_at_a_time_{label}"] = {"solves_per_s": len(ok) / makespan, "p50_s": float(np.percentile(lat, 50)),
                                             "p99_s": float(np.percentile(lat, 99)), "makespan_s": makespan}
        leg["speedup_vs_one_at_a_time"] = rep.solves_per_s / leg["one_at_a_time_solve"]["solves_per_s"]
        out["rates"][str(rate)] = leg
        print(json.dumps({"rate_done": rate, **leg}), flush=True)
    print(json.dumps(out))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    # 1000 C2 steps are about 30 ms of work: a device coming out of idle
    # (after the autotuning's host work and the state upload) clocks up over
    # 20-30 ms, and 400 steps timed right after 20 warmup steps read 3-4 %
    # slower than the same steps on a warm device
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--cpu-steps", type=int, default=480, help="oracle steps timed for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the full run to T=2 (for profiling)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-stream leg")
    ap.add_argument("--no-tgv", action="store_true", help="skip the 3D Taylor-Green leg")
    ap.add_argument("--no-advection", action="store_true", help="skip the 3D advection (C1) leg")
    ap.add_argument("--no-large", action="store_true", help="skip the 512^2 vortex leg")
    ap.add_argument("--ensemble", type=int, default=512, help="instances of the batched ensemble")
    ap.add_argument("--solo-sample", type=int, default=32, help="ensemble instances also solved one at a time")
    ap.add_argument("--requests", type=int, default=16, help="requests of the serving stream")
    ap.add_argument("--rate", type=float, default=8.0, help="serving arrival rate, requests/s")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the serving stream (--requests of them) at each rate of --rates")
    ap.add_argument("--rates", type=lambda v: [float(x) for x in v.split(",")], default=[4.0, 16.0],
                    help="arrival rates for --serve-only, requests/s, comma separated")
    ap.add_argument("--profile-batch", type=int, default=0, metavar="STEPS",
                    help="only run the batched ensemble for STEPS steps per instance, both forms, kernels "
                         "launched without graphs (for a kernel tracer)")
    ap.add_argument("--profile-large", type=int, default=0, metavar="STEPS",
                    help="only the 512^2 vortex leg, with STEPS more steps per form (for the kernel tracer)")
    ap.add_argument("--profile-limited", type=int, default=0, metavar="STEPS",
                    help="only a limited Sod run of STEPS steps per form (for the kernel tracer)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 1 GiB copy used to calibrate the HBM byte counters")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (got --gpus {args.gpus}); the multi-GPU halo "
              "design is documented but not built", file=sys.stderr)
        return 2

    import mojoxm_amd as M

    if args.profile_batch:
        return profile_batch(M, args)
    if args.profile_limited:
        return profile_limited(M, args)
    if args.profile_large:
        print(json.dumps(large_vortex_leg(M, args, nsteps=args.profile_large)))
        return 0
    if args.serve_only:
        return serve_study(M, args)

    pb = M.problems.euler_vortex_2d()
    K_el = pb.n[0] * pb.n[1] * 2
    mesh = M.LocalMesh2D(*pb.n, *pb.L, P=pb.P)
    q0 = pb.ic(mesh.elem_node_xyz)
    # both forms start from the same FP32-representable state
    q0 = q0.astype(np.float32).astype(np.float64)
    exact = pb.exact(mesh.elem_node_xyz, pb.t_final)

    forms = {"fp32_mfma": (M.PREC_FP32_MFMA, 4), "fp64": (M.PREC_FP64, 8)}
    res, finals = {}, {}
    for name, (prec, word) in forms.items():
        solver = M.Solver(pb.n, pb.L, pb.physics, P=pb.P, precision=prec)
        U = solver.dof_count()  # K*NP*NC DOF per stage
        # launch configuration: every occupancy target timed on this mesh
        solver.set_state(q0)
        cfg, cfg_us = solver.autotune(pb.dt, 30)  # leaves the state at q0
        tile = tile_report(solver)
        # ---- timed region: K steps with the state resident in HBM --------
        # (no upload here: host work between the autotuner's warm steps and
        # the timed ones lets the device idle and clock down)
        solver.prepare_steps(pb.dt, args.steps)  # graph construction, no steps
        # the warmup steps last, so the device is not idle when timing starts
        solver.step_ssprk3(pb.dt, args.warmup)
        solver.synchronize()
        t0 = time.perf_counter()
        solver.step_ssprk3(pb.dt, args.steps)
        solver.synchronize()
        wall = time.perf_counter() - t0
        # ---- stage-kernel duration from HIP events on the solver's stream
        solver.set_state(q0)
        avg_us = solver.time_stage_kernel(pb.dt, 50)
        bytes_per_launch = 8.0 / 3.0 * U * word  # 8 state touches per step of U values, per stage
        achieved = bytes_per_launch / (avg_us * 1e-6) / 1e9
        r = {"value": 3.0 * U * args.steps / wall, "ms_per_step": wall / args.steps * 1e3,
             "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(name),
                          "kernel_avg_us": avg_us, "algorithmic_bytes_per_launch": bytes_per_launch},
             "launch_config": {"chosen": cfg, "waves_per_simd": CFG_WAVES[name][cfg],
                               "us_per_config": dict(zip(CFG_WAVES[name], cfg_us)),
                               "tile": tile}}
        # ---- accuracy of the full run to T=2 ------------------------------
        if not args.no_accuracy:
            solver.set_state(q0)
            solver.step_ssprk3(pb.dt, pb.nsteps)
            finals[name] = solver.get_state()
            r["l2_vs_exact_T2"] = rel_l2(finals[name], exact)
            r["l2_vs_exact_T2_density"] = rel_l2(finals[name][..., 0], exact[..., 0])
        res[name] = (r, solver)

    # FP32 form against the FP64 form at T=2 (the FP64 form is tied to the
    # oracle at 1e-12 on the CPU sample below)
    if finals:
        d = finals["fp32_mfma"] - finals["fp64"]
        res["fp32_mfma"][0]["l2_vs_fp64_T2"] = rel_l2(finals["fp32_mfma"], finals["fp64"])
        res["fp32_mfma"][0]["linf_vs_fp64_T2"] = float(np.abs(d).max() / np.abs(finals["fp64"]).max())

    # ---- CPU baseline: the oracle on a bounded sample of the same run ----
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O

        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        O.set_threads(threads)
        om = O.Mesh(2, pb.P, pb.n, pb.L)
        n_cpu = args.cpu_steps
        c0 = time.perf_counter()
        q_cpu = om.step(pb.physics.phys_id, pb.physics.params(), q0, pb.dt, n_cpu)
        c_wall = time.perf_counter() - c0
        for name, (r, solver) in res.items():
            solver.set_state(q0)
            solver.step_ssprk3(pb.dt, n_cpu)
            q_gpu = solver.get_state()
            r["l2_vs_oracle_sample"] = rel_l2(q_gpu, q_cpu)
            r["linf_vs_oracle_sample"] = float(np.abs(q_gpu - q_cpu).max() / np.abs(q_cpu).max())
        U = res["fp64"][1].dof_count()
        # the same oracle on one core, over a shorter sample of the run
        n_one = max(1, n_cpu // 24)
        O.set_threads(1)
        c1 = time.perf_counter()
        om.step(pb.physics.phys_id, pb.physics.params(), q0, pb.dt, n_one)
        one_wall = time.perf_counter() - c1
        O.set_threads(threads)
        cpu = {"value": 3.0 * U * n_cpu / c_wall, "unit": "DOF-stage-updates/s", "cores": threads,
               "kind": "port",
               "sample": f"{n_cpu} of {pb.nsteps} SSPRK3 steps of the same 128^2 P=3 vortex run "
                         f"({c_wall:.1f} s on {threads} OpenMP threads)",
               "one_core": {"value": 3.0 * U * n_one / one_wall, "cores": 1,
                            "sample": f"{n_one} steps of the same run ({one_wall:.1f} s on one thread)"}}
    U = res["fp64"][1].dof_count()
    for _, solver in res.values():
        solver.close()

    advection = advection_leg(M, args) if not args.no_advection else None
    tgv = taylor_green_leg(M, args) if not args.no_tgv else None
    large = large_vortex_leg(M, args) if not args.no_large else None
    batched = batched_leg(M, args) if not args.no_batch else None
    serving = serving_leg(M, args) if not args.no_serve else None

    r32, r64 = res["fp32_mfma"][0], res["fp64"][0]
    fp64 = {k: v for k, v in r64.items() if k not in ("roofline",)}
    out = {
        "metric": METRIC,
        "value": r32["value"],
        "unit": "DOF-stage-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "euler_vortex_2d: 128x128x2 tris, P=3, NC=5, HLLEC, CFL 0.15, T=2",
            "kernel": "FP32-MFMA fused stage kernel (FP32 state and fluxes, volume+LIFT on "
                      "V_MFMA_F32_16X16X4F32, FP64 geometry, combine and positivity)",
            "elements": K_el, "dof_per_stage": U, "dt": pb.dt,
            "steps_to_T": pb.nsteps,
            "fp32_mfma": {k: v for k, v in r32.items() if k not in ("roofline", "value", "ms_per_step")},
            "fp64": fp64,
            "speedup_fp32_mfma_vs_fp64": r32["value"] / r64["value"],
        },
        "advection": advection,
        "taylor_green": tgv,
        "vortex_512": large,
        "batched": batched,
        "serving": serving,
        "roofline": r32["roofline"],
        "roofline_fp64": r64["roofline"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
