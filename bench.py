#!/usr/bin/env python3
"""Particle-steps/s of batched mixed-precision FLIP serving vs the plain one-at-a-time FP32
path on one MI355X, with p50/p99 request latency.

Default workload (`--workload serve`): a stream of independent liquid simulations.  Request i
is a cubic grid n^3 on the unit cube with n drawn from {32, 48, 64, 96, 128}, one liquid box
or (even odds) sphere of random place and size in the interior with a random initial
velocity, 4-8 particles per
cell (fp32 SoA state), dt = f32(1/120), PIC/FLIP 0.95, seeded with seed = i, and `--steps`
steps.  Requests arrive as a Poisson process at a server with a 200 GB HBM admission budget.
The arrival rate is `--rate`, or by default the steady-state service rate of a fixed batch of
the stream's first `--warmup` (64) requests, run untimed beforehand, in requests per second
of this stream's mean work: the load a server holding about that many simulations at once
sustains (the batch grows past it when arrivals bunch up and drains back when they thin).
Each simulation's pressure solve runs to a relative FP64 residual of 1e-6 (multigrid CG in
FP32 inside FP64 refinement).  `value` = sum over requests of particles x steps, divided by
the time from the first submission to the last completion (whole job, inputs created on the
device at admission).

`plain` runs a sample of the same requests one at a time through the single-simulation path
with the FP32 pressure CG at a fixed iteration count (the reference's default 80), creation
included, and `ratio_vs_plain` = value / plain.

`roofline` is for the kernel with the largest share of time in a profiled batched step of a
fixed batch of stream requests (HIP events on the batch stream, launches priced by their
algorithmic bytes over the simulations each covers).  `cpu_baseline` times the C oracle
(single thread) on a bounded sample of the stream's requests.

`--workload dam` measures the single-instance 128^3 dam break on the plain path instead;
`--workload batch` only the fixed batch used for the roofline (for profiling runs);
`--workload mixed256` the 256^3 two-block mixed-precision configuration (~58M particles);
`--workload obstacles` the 192x96x192 obstacle-heavy configuration (solid spheres, 19.4M
particles, 200 steps, mixed-precision solve); `--workload frame` 128^3 dam-break frames of 5
steps plus one surface-nets pass; `--workload reference64` the reference's own 64^3 dam break
in reference mode (RBGS 80) with the CFL cap, beside the CPU oracle on the same configuration;
`--workload agents` the agent neighbour grid of a 1M-agent flock.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
METRIC = "particle-steps/s batched vs plain at ≤1e-6 rel residual; p50/p99 sim latency"


def plain_algorithmic_bytes(name: str, P: int, F: int, C: int, Cf: int) -> float:
    """Minimal bytes one single-sim launch must move: every logical array once, at its element
    width (fp32/int32 4 B, cell type 1 B, stencil 2 B, particle x+v 24 B)."""
    table = {
        "clear_grid": 8 * F + 4 * C,
        "bin_count": 12 * P + 8 * P + 4 * C,
        "scan_offsets": 8 * C,
        "scatter": 36 * P + 28 * P + 4 * C,
        "p2g": 24 * P + 8 * F,
        "faces": 16 * F,
        "cells": 4 * C + 4 * F + 7 * C + 4 * (C - Cf),
        "cg_start": 2 * C + 12 * Cf + 2 * C + 16 * Cf,
        "cg_iter": 2 * C + 40 * Cf,
        "project": 8 * F + 5 * C,
        "g2p": 28 * P + 8 * F + 28 * P,
    }
    return float(table.get(name, 0))


def bench_dam(args, pxl):
    from tests.scenes import FPS, dam_break
    fluid = dam_break(128, iters=80)
    sim = pxl.FluidState(fluid, FPS, seed=0, solver="cg", cg_iterations=80)
    P = sim.particle_count()
    info = sim.info
    for _ in range(args.warmup):
        sim.step(1)
    tuned = sim.autotune(2)
    t0 = time.perf_counter()
    sim.step(args.steps)
    elapsed = time.perf_counter() - t0
    prof = sim.profile_steps(2)
    Cf = int((sim.read_field("type") == 1).sum())
    F, Cc = info["n_faces"], info["n_cells"]
    kernels = []
    for name, launches, total_ms in prof:
        avg_s = total_ms / 1e3 / launches
        b = plain_algorithmic_bytes(name, P, F, Cc, Cf)
        kernels.append({"kernel": name, "launches_per_step": launches // 2, "avg_us": avg_s * 1e6,
                        "alg_bytes": b, "achieved_GBs": (b / avg_s / 1e9) if b else None})
    dom = kernels[0]
    ach = dom["achieved_GBs"] or 0.0
    sim.close()
    traffic = measured_traffic_pmc(dom["kernel"], ["--probe-kind", "dam"]) if args.traffic else None
    if traffic and dom["alg_bytes"]:
        traffic["traffic_over_alg"] = traffic["hbm_bytes"] / dom["alg_bytes"]
    cpu = cpu_dam(args, fluid) if args.cpu_seconds > 0 else None
    return {
        "metric": METRIC, "value": P * args.steps / elapsed, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded dam break, reference LCG seeding)",
        "config": {"workload": "dam break 128^3, single instance, plain FP32 Jacobi-PCG, 80 fixed iterations",
                   "particles": P, "autotune": tuned},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": ach / PEAK_HBM_GBS, "traffic": traffic["hbm_bytes"] if traffic else None,
                     "traffic_detail": traffic,
                     "kernel": dom["kernel"], "avg_us": dom["avg_us"], "alg_bytes_per_launch": dom["alg_bytes"]},
        "cpu_baseline": cpu, "kernels": kernels,
    }


def surface_algorithmic_bytes(stage, P, nodes, cells, touched, V, T):
    """Minimal HBM bytes of one surface stage: positions 12 B per particle, 4 B per node or
    cell word read or written once, an atomic-add target (read + write) 8 B per touched node,
    32 B per vertex, 12 B per triangle."""
    return {"s_clear": 4 * nodes, "s_splat": 12 * P + 8 * touched, "s_phi": 8 * nodes,
            "s_vflag": 4 * nodes + 4 * cells, "s_vscan": 12 * cells, "s_verts": 4 * nodes + 8 * cells + 32 * V,
            "s_qcount": 4 * nodes + 4 * nodes, "s_qscan": 12 * nodes,
            "s_quads": 8 * nodes + 12 * T}.get(stage, 0)


# kernel symbols of the surface stages (rocprofv3 names)
SURF_KERNELS = {"s_splat": "k_s_splat_lds", "s_phi": "k_s_phi", "s_vflag": "k_s_vflag", "s_verts": "k_s_verts",
                "s_qcount": "k_s_qcount", "s_quads": "k_s_quads"}


def bench_frame(args, pxl):
    """Frames of the 128^3 dam break: `substeps` (5) plain FP32 CG-80 steps plus one surface-nets
    pass (refine 2: a 256^3 surface grid) per frame, the reference's per-frame work
    (fluid.rs:265-330).  value = particle-steps/s with the meshing included; the roofline is
    the surface pass's dominant stage."""
    import numpy as np
    from tests.scenes import FPS, dam_break
    fluid = dam_break(128, iters=80)
    sim = pxl.FluidState(fluid, FPS, seed=0, solver="cg", cg_iterations=80)
    P = sim.particle_count()
    sub = fluid["substeps"]
    for _ in range(args.warmup):
        sim.step(sub)
        sim.surface()
    frames = max(1, args.steps // sub)
    cap = sim.reservation()[0]
    vbuf, ibuf = pxl.DeviceBuffer(32 * cap), pxl.DeviceBuffer(12 * cap)  # the renderer's buffers
    t_mesh = 0.0
    t0 = time.perf_counter()
    for _ in range(frames):
        sim.step(sub)
        a = time.perf_counter()
        info = sim.surface_device(vbuf.ptr, 0, ibuf.ptr, 0)
        t_mesh += time.perf_counter() - a
    elapsed = time.perf_counter() - t0
    vbuf.free()
    ibuf.free()
    prof = sim.profile_surface(3)
    dens = sim.read_field("density")
    sx, sy, sz = info["sx"], info["sy"], info["sz"]
    nodes, cells = (sx + 1) * (sy + 1) * (sz + 1), sx * sy * sz
    touched = int(np.count_nonzero(dens))
    V, T = info["vertices"], info["triangles"]
    stages = []
    for name, launches, total_ms in prof:
        avg_s = total_ms / 1e3 / launches
        b = surface_algorithmic_bytes(name, P, nodes, cells, touched, V, T)
        stages.append({"kernel": name, "avg_us": avg_s * 1e6, "alg_bytes": b,
                       "achieved_GBs": b / avg_s / 1e9 if b else None})
    dom = stages[0]
    mesh_ms = sum(st["avg_us"] for st in stages) / 1e3
    sim.close()
    traffic = measured_traffic_pmc(dom["kernel"], ["--probe-kind", "frame"]) if args.traffic else None
    if traffic and dom["alg_bytes"]:
        traffic["traffic_over_alg"] = traffic["hbm_bytes"] / dom["alg_bytes"]
    return {
        "metric": METRIC, "value": P * frames * sub / elapsed, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": frames * sub, "warmup": args.warmup, "ms_per_step": elapsed / (frames * sub) * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded dam break, reference LCG seeding)",
        "config": {"workload": f"dam break 128^3 frames: {sub} plain FP32 CG-80 steps + surface nets (256^3 surface grid) "
                               f"per frame, {frames} frames",
                   "particles": P, "surface_ms_per_frame_wall": t_mesh / frames * 1e3,
                   "surface_ms_per_frame_kernels": mesh_ms, "vertices": V, "triangles": T,
                   "surface_nodes": nodes, "touched_nodes": touched, "reservation": cap},
        "roofline": {"bound": "hbm", "achieved": dom["achieved_GBs"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": (dom["achieved_GBs"] or 0.0) / PEAK_HBM_GBS,
                     "traffic": traffic["hbm_bytes"] if traffic else None, "traffic_detail": traffic,
                     "kernel": dom["kernel"], "avg_us": dom["avg_us"], "alg_bytes_per_launch": dom["alg_bytes"]},
        "cpu_baseline": None, "kernels": stages,
    }


def bench_reference64(args, pxl):
    """The reference's own configuration: the 64^3 dam break (458,800 particles) in reference
    mode (red-black Gauss-Seidel, 80 iterations per step, bit-exact against the oracle) with
    the CFL cap at one cell per substep, --steps (100) steps; the CPU baseline is the oracle
    on the same configuration (first steps, about --cpu-seconds)."""
    from tests.scenes import FPS, dam_break
    fluid = dict(dam_break(64, iters=80), cfl=1.0)
    sim = pxl.FluidState(fluid, FPS, seed=0)
    P = sim.particle_count()
    ks = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.step(1)
        ks.append(sim.last_substeps())
    elapsed = time.perf_counter() - t0
    prof = sim.profile_steps(2)
    sim.close()
    cpu = None
    if args.cpu_seconds > 0:
        import oracle
        from tests.scenes import oracle_sim
        osim = oracle_sim(oracle, fluid)
        n, spent = 0, 0.0
        while n == 0 or (spent < args.cpu_seconds and n < args.steps):
            a = time.perf_counter()
            osim.step_rbgs_capped(80, fluid["cfl"])
            spent += time.perf_counter() - a
            n += 1
        cpu = {"value": P * n / spent, "unit": "particle-steps/s", "cores": 1, "kind": "port",
               "sample": f"first {n} capped RBGS-80 steps of the same 64^3 dam break, C oracle, one thread"}
    return {
        "metric": METRIC, "value": P * args.steps / elapsed, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 0, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded dam break, reference LCG seeding)",
        "config": {"workload": f"dam break 64^3, single instance, reference mode (RBGS 80), CFL cap 1 cell per "
                               f"substep, {args.steps} steps from seeding",
                   "particles": P, "substeps_histogram": {str(k): ks.count(k) for k in sorted(set(ks))}},
        "roofline": None, "cpu_baseline": cpu,
        "phases": [{"phase": name, "launches_per_step": launches // 2, "ms_per_step": total_ms / 2}
                   for name, launches, total_ms in prof],
    }


def flock_agents(n: int, seed: int = 0):
    """n agents (64-byte records) in 16 Gaussian clumps of random size and spread, like the
    reference's 1M-starling flock example (examples/programs/flock.slang)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-120.0, 120.0, (16, 3))
    spreads = rng.uniform(4.0, 25.0, 16)
    which = rng.integers(0, 16, n)
    a = np.zeros((n, 16), np.float32)
    a[:, 0:3] = (centres[which] + rng.normal(0.0, 1.0, (n, 3)) * spreads[which, None]).astype(np.float32)
    a[:, 4:7] = rng.normal(0.0, 3.0, (n, 3)).astype(np.float32)
    return a


def bench_agents(args, pxl):
    """The agent neighbour grid (shaders/sim.slang:27-104) of the reference's flock example:
    1M agents, cell 1.6, table 2^21, built --steps times from device-resident agents; value =
    agents binned per second."""
    import numpy as np
    n, cell = 1_000_000, 1.6
    a = flock_agents(n)
    t = pxl.AgentGrid.table_size(n)
    g = pxl.AgentGrid(n)
    bufs = [pxl.DeviceBuffer(b) for b in (a.nbytes, 4 * t, 4 * t, 4 * n)]
    bufs[0].upload(a)
    for _ in range(max(args.warmup, 1)):
        g.build_device(bufs[0].ptr, n, cell, t, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr)
    ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ms.append(g.build_device(bufs[0].ptr, n, cell, t, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr))
    elapsed = time.perf_counter() - t0
    cnt = bufs[2].download(np.zeros(t, np.uint32))
    for b in bufs:
        b.free()
    g.close()
    build_ms = float(np.mean(ms))
    alg = 12 * n + 4 * n + 8 * t  # positions read, ids written, counts and offsets written
    ach = alg / (build_ms / 1e3) / 1e9
    return {
        "metric": "agent neighbour-grid builds", "value": n * args.steps / elapsed, "unit": "agents/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (1M agents in 16 Gaussian clumps)",
        "config": {"workload": f"agent neighbour grid: {n} agents, cell {cell}, table {t}, device-resident",
                   "build_ms_events": build_ms, "occupied_cells": int(np.count_nonzero(cnt)), "max_per_cell": int(cnt.max())},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach / PEAK_HBM_GBS,
                     "traffic": None, "kernel": "grid build (all launches)", "avg_us": build_ms * 1e3,
                     "alg_bytes_per_launch": alg},
        "cpu_baseline": None,
    }


def bench_single_mixed(args, pxl, fluid, data, workload, max_cells=0):
    """One large simulation with the MIXED solve (FP32 multigrid CG inside FP64 refinement) to
    1e-6 every step: particle-steps/s over --steps steps after --warmup, the FP64 relative
    residual and refinement passes of every step, and the per-phase times of two profiled
    steps."""
    from tests.scenes import FPS
    sim = pxl.FluidState(fluid, FPS, seed=0, solver="mixed", tolerance=1e-6, max_cells=max_cells,
                         inner_max=args.inner_max)
    P = sim.particle_count()
    for _ in range(args.warmup):
        sim.step(1)
    res, outer = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.step(1)
        st = sim.last_solve()
        res.append(st["rel_residual"])
        outer.append(st["outer_iterations"])
    elapsed = time.perf_counter() - t0
    prof = sim.profile_steps(2)
    sim.close()
    return {
        "metric": METRIC, "value": P * args.steps / elapsed, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": data,
        "config": {"workload": workload, "particles": P, "max_rel_residual": max(res),
                   "pass_length": f"adaptive, 2..{args.inner_max}" if args.inner_max else "fixed (9, then 2, 4, ...)",
                   "outer_passes": {str(k): outer.count(k) for k in sorted(set(outer))},
                   "residual_per_step": [float(f"{r:.3g}") for r in res]},
        "roofline": None, "cpu_baseline": None,
        "phases": [{"phase": name, "launches_per_step": launches // 2, "ms_per_step": total_ms / 2}
                   for name, launches, total_ms in prof],
    }


def bench_mixed256(args, pxl):
    """The mixed-precision configuration: 256^3 grid (above the reference's 16M-cell limit, an
    extension enabled with max_cells), two liquid blocks colliding, ~58M particles."""
    from tests.scenes import two_blocks
    return bench_single_mixed(args, pxl, two_blocks(256), "synthetic (two colliding liquid blocks, reference LCG seeding)",
                              "256^3 grid, two blocks at 8 particles/cell colliding at 1.5 m/s, single instance, "
                              "MIXED solve to 1e-6 every step", max_cells=1 << 25)


def bench_obstacles(args, pxl):
    """The obstacle-heavy configuration: 192 x 96 x 192 grid on [0,2] x [0,1] x [0,2], solid
    spheres (an extension) of radius 2-8 cells at seeded random places covering ~5% of the
    volume, a 19.4M-particle liquid slab falling onto them."""
    from tests.scenes import obstacle_slab
    f = obstacle_slab(192)
    return bench_single_mixed(args, pxl, f, "synthetic (seeded solid spheres, liquid slab, reference LCG seeding)",
                              f"192x96x192 grid, {len(f['obstacles'])} solid spheres of radius 2-8 cells (~5% of the "
                              "volume), liquid slab y in [0.25, 1] at 8 particles/cell, single instance, MIXED solve "
                              "to 1e-6 every step")


def cpu_dam(args, fluid):
    import oracle
    from tests.scenes import oracle_sim
    osim = oracle_sim(oracle, fluid)
    t0 = time.perf_counter()
    n = 0
    while n == 0 or (time.perf_counter() - t0 < args.cpu_seconds and n < 4):
        osim.step_cg(80)
        n += 1
    dt = time.perf_counter() - t0
    return {"value": osim.P * n / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"first {n} steps of the 128^3 dam break ({osim.P} particles), FP32 CG 80, C oracle, one thread"}


def _stream(pxl, srv, reqs, arrivals, steps):
    """Submits reqs[i] (seed i) at arrivals[i] seconds after the start and waits for every one
    to complete; returns (wall seconds from before the first submission until the last
    completion was seen, request ids, server stats).  The release of the last simulations'
    memory (drain) follows outside that time."""
    t0 = time.perf_counter()
    ids = []
    for i, f in enumerate(reqs):
        delay = arrivals[i] - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(f, FPS_DEFAULT, seed=i, steps=steps))
    for i in ids:
        srv.wait(i)
    wall = time.perf_counter() - t0
    srv.drain()
    return wall, ids, srv.stats()


FPS_DEFAULT = 24.0
REQUEST_STEPS = 60  # steps per serving request (the serving configuration)


def bench_serve(args, pxl):
    """The serving configuration: 512 requests of 60 steps each.

    value: capacity -- the whole stream submitted at once (offered load above what the server
    can take, so the device never waits for arrivals), particle-steps over the time from the
    first submission to the last completion; the timed region is that run, and `steps` is the
    number of batched steps the server ran in it (ms_per_step x steps = the timed region).
    Latency: the same stream again with Poisson arrivals at --load (0.7) of that capacity, in
    requests per second of the stream's mean work; p50/p99 are submit -> completion there.
    small_batch: the first --small-requests requests at once with at most --small-batch in the
    batch (capacity where the batched step's fixed cost is not amortised).
    plain: every --plain-every-th request alone on the single-simulation FP32 CG-80 path,
    creation included.  roofline: two event-timed batched steps of the batch the capacity run
    held at its start (the requests it admitted at once)."""
    import numpy as np
    from tests.scenes import FPS, random_request

    sizes = tuple(int(v) for v in args.sizes.split(","))
    reqs = [random_request(i, sizes=sizes) for i in range(args.requests)]
    opts = dict(budget_bytes=args.budget, max_sims=args.max_sims, inner_iterations=args.inner_iterations)
    infos = [pxl.check(f, FPS, seed=i, solver="mixed") for i, f in enumerate(reqs)]
    P = np.array([inf["n_particles"] for inf in infos], dtype=np.float64)
    nbytes = np.array([inf["device_bytes"] for inf in infos], dtype=np.float64)
    work = P * REQUEST_STEPS

    # warm-up (untimed): device initialisation, kernel loading and graph capture on a few
    # requests of the same distribution, --warmup batched steps
    with pxl.Batch([random_request(100000 + i, sizes=sizes) for i in range(8)], FPS,
                   seeds=list(range(8))) as bat:
        bat.step(max(args.warmup, 1))

    # capacity (timed): every request submitted at once; every --profile-every-th batched step
    # runs its launches eagerly, each timed by HIP events (the roofline's kernel durations)
    with pxl.Server(profile_every=args.profile_every, **opts) as srv:
        wall, ids, st = _stream(pxl, srv, reqs, np.zeros(len(reqs)), REQUEST_STEPS)
        per = [srv.poll(i) for i in ids]
        kstats = srv.kernel_stats()
    failed = [p for p in per if p["status"] != "done"]
    value = st["particle_steps"] / wall
    # the requests the capacity run admitted at once: first in, first out within max_sims and
    # the budget
    first = int(min(args.max_sims, np.searchsorted(np.cumsum(nbytes), args.budget, side="right")))

    # latency: Poisson arrivals at --load of the measured capacity
    rate = args.load * value / work.mean()
    rng = np.random.default_rng(12345)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(reqs)))
    arrivals -= arrivals[0]
    with pxl.Server(**opts) as srv:
        lwall, lids, lst = _stream(pxl, srv, reqs, arrivals, REQUEST_STEPS)
        lper = [srv.poll(i) for i in lids]
    failed += [p for p in lper if p["status"] != "done"]

    # small batches: the first --small-requests requests at once, at most --small-batch in the
    # batch (the batched step's fixed cost, which a deep batch amortises)
    small = None
    if args.small_batch > 0 and args.small_requests > 0:
        with pxl.Server(**dict(opts, max_sims=args.small_batch)) as srv:
            swall, sids, sst = _stream(pxl, srv, reqs[:args.small_requests], np.zeros(args.small_requests),
                                       REQUEST_STEPS)
            sper = [srv.poll(i) for i in sids]
        failed += [p for p in sper if p["status"] != "done"]
        small = {"requests": args.small_requests, "max_batch": args.small_batch, "offered": "all requests at t = 0",
                 "value": sst["particle_steps"] / swall, "wall_s": swall, "mean_batch": sst["mean_batch"],
                 "batch_steps": sst["batch_steps"], "ms_per_step": 1e3 * swall / max(sst["batch_steps"], 1),
                 "busy_fraction": sst["busy_s"] / swall, "p50_ms": sst["p50_ms"], "p99_ms": sst["p99_ms"],
                 "max_rel_residual": max(p["rel_residual"] for p in sper)}

    # plain: a sample of the same requests, one at a time, FP32 CG with 80 iterations
    sample = list(range(0, len(reqs), max(1, args.plain_every)))
    pw, pt, pt_steps = 0.0, 0.0, 0.0
    for i in sample:
        a = time.perf_counter()
        sim = pxl.FluidState(reqs[i], FPS, seed=i, solver="cg", cg_iterations=80)
        b = time.perf_counter()
        sim.step(REQUEST_STEPS)
        c = time.perf_counter()
        pw += sim.particle_count() * REQUEST_STEPS
        sim.close()
        d = time.perf_counter()
        pt += d - a
        pt_steps += c - b
    plain = {"value": pw / pt, "steps_only": pw / pt_steps, "requests": len(sample),
             "path": "one simulation at a time, FP32 Jacobi-PCG 80 fixed iterations, hipGraph per step, "
                     "creation included (steps_only: step time alone)"}

    roofline, kernels = roofline_of(kstats)
    profiled = sum(k["launches"] for k in kstats if k["kernel"] == "g2p")
    roofline["measured_on"] = (f"HIP events of every launch of {profiled} of the capacity run's {st['batch_steps']} "
                               f"batched steps (every {args.profile_every}th step run eagerly)")
    if args.traffic:
        t = measured_traffic_pmc(roofline["kernel"], ["--profile-batch", str(first), "--sizes", args.sizes])
        if t:
            t["alg_bytes_per_launch"] = next((k["alg_bytes_per_launch"] for k in probe_kernels(pxl, reqs[:first])
                                              if k["kernel"] == roofline["kernel"]), None)
            if t["alg_bytes_per_launch"]:
                t["traffic_over_alg"] = t["hbm_bytes"] / t["alg_bytes_per_launch"]
            roofline["traffic"] = t["hbm_bytes"]
            roofline["traffic_detail"] = dict(t, per="dispatch, two event-timed batched steps of the same batch "
                                                     "under rocprofv3 --pmc (two passes)")

    cpu = cpu_serve(args, reqs) if args.cpu_seconds > 0 else None
    return {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": st["batch_steps"],
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / max(st["batch_steps"], 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (random liquid boxes or spheres, reference LCG seeding, seed = request index)",
        "config": {
            "workload": f"serving stream: {len(reqs)} requests, cubic grids n in {{{args.sizes}}}, one random "
                        f"liquid box or sphere (even odds) at 4-8 particles/cell, {REQUEST_STEPS} steps each, "
                        f"{args.budget / 1e9:.0f} GB admission budget, at most {args.max_sims} in the batch, "
                        f"MIXED solve to 1e-6",
            "steps_note": (f"steps = batched steps of the timed capacity run (--steps {args.steps} does not apply: "
                           f"each request is {REQUEST_STEPS} steps)"),
            "requests": len(reqs), "sphere_fills": sum("center" in f["fill"][0] for f in reqs),
            "particles_total": float(P.sum()),
            "capacity": {"offered": "all requests at t = 0", "wall_s": wall, "span_s": st["span_s"],
                         "mean_batch": st["mean_batch"], "batch_steps": st["batch_steps"],
                         "busy_fraction": st["busy_s"] / wall, "admit_s": st["admit_s"],
                         "graph_captures": st["graph_captures"], "capture_s": st["capture_s"],
                         "peak_bytes": st["peak_bytes"],
                         "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"]},
            "latency": {"arrivals": "Poisson", "load": args.load, "rate_per_s": rate,
                        "rate_rule": f"{args.load} x capacity / mean request work ({work.mean():.4g} particle-steps)",
                        "p50_ms": lst["p50_ms"], "p99_ms": lst["p99_ms"], "mean_ms": lst["mean_ms"],
                        "max_ms": lst["max_ms"], "mean_batch": lst["mean_batch"],
                        "throughput": lst["particle_steps"] / lwall, "wall_s": lwall,
                        "busy_fraction": lst["busy_s"] / lwall},
            "p50_ms": lst["p50_ms"], "p99_ms": lst["p99_ms"],
            "failed": len(failed), "max_rel_residual": max(p["rel_residual"] for p in per + lper),
            "plain": plain, "ratio_vs_plain": value / plain["value"],
            "small_batch": dict(small, ratio_vs_plain=small["value"] / plain["value"]) if small else None,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }


def probe_kernels(pxl, reqs):
    """Per-kernel rows of the probe's launches (the same batch, two event-timed steps), for the
    algorithmic bytes that go with its counters."""
    from tests.scenes import FPS
    with pxl.Batch(reqs, FPS, seeds=list(range(len(reqs)))) as bat:
        return roofline_of(bat.profile_steps(2))[1]


def roofline_of(prof):
    """(roofline of the kernel with the largest time share, per-kernel rows) from per-kernel
    totals [{"kernel", "launches", "total_ms", "alg_bytes"}]."""
    prof = sorted(prof, key=lambda k: -k["total_ms"])
    tot = sum(k["total_ms"] for k in prof)
    kernels = [{"kernel": k["kernel"], "launches": k["launches"], "avg_us": 1e3 * k["total_ms"] / k["launches"],
                "share": k["total_ms"] / tot, "alg_bytes_per_launch": k["alg_bytes"] / k["launches"],
                "achieved_GBs": (k["alg_bytes"] / (k["total_ms"] / 1e3) / 1e9) if k["alg_bytes"] else None}
               for k in prof]
    dom = kernels[0]
    ach = dom["achieved_GBs"] or 0.0
    roofline = {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach / PEAK_HBM_GBS,
                "traffic": None, "kernel": dom["kernel"], "avg_us": dom["avg_us"],
                "alg_bytes_per_launch": dom["alg_bytes_per_launch"]}
    return roofline, kernels


def batch_roofline(pxl, reqs, warm, timed, **opts):
    """A fixed batch of `reqs` (seed = index): `warm` untimed steps, `timed` timed steps, then
    two steps with every batched launch timed by HIP events and priced in algorithmic bytes.
    Returns (roofline of the kernel with the largest time share, per-kernel rows, particle-steps/s
    of the timed steps or None)."""
    from tests.scenes import FPS
    n = len(reqs)
    rate = None
    with pxl.Batch(reqs, FPS, seeds=list(range(n)), **opts) as bat:
        bat.step(warm)
        if timed:
            t0 = time.perf_counter()
            bat.step(timed)
            rate = sum(i["n_particles"] for i in bat.infos) * timed / (time.perf_counter() - t0)
        prof = bat.profile_steps(2)
    prof.sort(key=lambda k: -k["total_ms"])
    tot = sum(k["total_ms"] for k in prof)
    kernels = [{"kernel": k["kernel"], "launches": k["launches"], "avg_us": 1e3 * k["total_ms"] / k["launches"],
                "share": k["total_ms"] / tot, "alg_bytes_per_launch": k["alg_bytes"] / k["launches"],
                "achieved_GBs": (k["alg_bytes"] / (k["total_ms"] / 1e3) / 1e9) if k["alg_bytes"] else None}
               for k in prof]
    dom = kernels[0]
    ach = dom["achieved_GBs"] or 0.0
    roofline = {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach / PEAK_HBM_GBS,
                "traffic": None, "kernel": dom["kernel"],
                "avg_us": dom["avg_us"], "alg_bytes_per_launch": dom["alg_bytes_per_launch"],
                "measured_on": f"two profiled batched steps of the first {n} stream requests"}
    return roofline, kernels, rate


# HBM traffic from rocprofv3 counters.  Two passes (a pass may hold at most 4 TCC counters):
# EA read requests by size, then EA write requests by size and the DRAM-bound ones.  Bytes =
# 32 n32 + 64 n64 + 128 n128 (reads), 32 (n - n64) + 64 n64 (writes), each scaled by what the
# same formula reads for a device-to-device copy of known size (the factor is reported).
PMC_PASSES = (("TCC_EA0_RDREQ_sum", "TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum"),
              ("TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum", "TCC_EA0_RDREQ_DRAM_sum", "TCC_EA0_WRREQ_DRAM_sum"))
CALIB_BYTES = 256 << 20
COPY_KERNEL = "__amd_rocclr_copyBuffer"


def kernel_symbol(name: str) -> list:
    """Symbols the launch of phase `name` may carry: batched (kb_sweep for mg_sweep, kb_g2p_tile
    for g2p, ...), single-simulation (k_cg_iter, k_g2p_tile, ...) or a surface stage (k_s_splat_lds)."""
    base = name[3:] if name.startswith("mg_") and name not in ("mg_tail", "mg_types", "mg_sten", "mg_scan", "mg_compact") else name
    out = [f"kb_{name}", f"kb_{base}", f"kb_{name}_tile", f"k_{name}", f"k_{name}_tile", f"k_{name}_lds"]
    return out + ([SURF_KERNELS[name]] if name in SURF_KERNELS else [])


def pmc_counters(csv_dir: str) -> dict:
    """{kernel symbol: [{counter: value} per dispatch]} from a rocprofv3 counter-collection CSV
    (symbols without namespaces or arguments: kb_sweep, __amd_rocclr_copyBuffer, ...)."""
    import csv
    import glob
    out = {}
    for path in glob.glob(os.path.join(csv_dir, "**", "*counter_collection.csv"), recursive=True):
        per = {}
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                name = row["Kernel_Name"].replace("(anonymous namespace)", "anon").split("(")[0]
                name = name.split("<")[0].split("::")[-1].split(" ")[-1]
                per.setdefault((name, row["Dispatch_Id"]), {})[row["Counter_Name"]] = float(row["Counter_Value"])
        for (name, _), ctr in per.items():
            out.setdefault(name, []).append(ctr)
    return out


def measured_traffic_pmc(kernel: str, probe_args: list, timeout: float = 240.0):
    """HBM bytes per dispatch of the batched `kernel` in the probe workload, from two rocprofv3
    --pmc passes run as child processes, with the copy calibration and the per-size split.
    Returns a dict, or None when rocprofv3 is unavailable or a pass fails."""
    import shutil
    import subprocess
    import tempfile

    import numpy as np
    prof = shutil.which("rocprofv3")
    if not prof:
        return None
    res, dispatches = {}, 0
    env = dict(os.environ, TMPDIR="/tmp")
    for counters in PMC_PASSES:
        d = tempfile.mkdtemp(prefix="pxl_pmc_", dir="/tmp")
        cmd = [prof, "--pmc", *counters, "--output-format", "csv", "-d", d, "-o", "run", "--",
               sys.executable, os.path.join(ROOT, "bench.py"), "--workload", "probe", *probe_args]
        try:
            r = subprocess.run(cmd, cwd=ROOT, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                               timeout=timeout, text=True)
        except subprocess.TimeoutExpired:
            shutil.rmtree(d, ignore_errors=True)
            return None
        data = pmc_counters(d)
        shutil.rmtree(d, ignore_errors=True)
        ks = next((data[s] for s in kernel_symbol(kernel) if s in data), None)
        copies = data.get(COPY_KERNEL)
        if r.returncode != 0 or not ks or not copies:
            return None
        big = max(copies, key=lambda c: c.get(counters[0], 0.0))  # the calibration copy
        dispatches = len(ks)
        for name in counters:
            res[name] = (float(np.mean([c.get(name, 0.0) for c in ks])), big.get(name, 0.0))

    def rd(i):
        n32, n64, n128 = (res[f"TCC_EA0_RDREQ_{s}_sum"][i] for s in ("32B", "64B", "128B"))
        return 32 * n32 + 64 * n64 + 128 * n128

    def wr(i):
        n, n64 = res["TCC_EA0_WRREQ_sum"][i], res["TCC_EA0_WRREQ_64B_sum"][i]
        return 32 * (n - n64) + 64 * n64
    f_rd = CALIB_BYTES / rd(1) if rd(1) else 1.0
    f_wr = CALIB_BYTES / wr(1) if wr(1) else 1.0
    read, write = rd(0) * f_rd, wr(0) * f_wr
    return {"hbm_bytes": read + write, "read_bytes": read, "write_bytes": write, "dispatches": dispatches,
            "read_factor": f_rd, "write_factor": f_wr,
            "dram_read_fraction": res["TCC_EA0_RDREQ_DRAM_sum"][0] / max(res["TCC_EA0_RDREQ_sum"][0], 1.0),
            "dram_write_fraction": res["TCC_EA0_WRREQ_DRAM_sum"][0] / max(res["TCC_EA0_WRREQ_sum"][0], 1.0),
            "counters": {k: v[0] for k, v in res.items()},
            "copy_counters": {k: v[1] for k, v in res.items()}}


def bench_probe(args, pxl):
    """Counter probe, run under rocprofv3 --pmc by measured_traffic_pmc: a device-to-device copy
    of CALIB_BYTES (the calibration), then the launches the roofline of --probe-kind prices:
    serve -- --probe-steps event-timed batched steps of the fixed batch of the first
    --profile-batch stream requests; dam -- two steps of the 128^3 dam break (FP32 CG 80) after
    one; frame -- the 128^3 dam break's surface pass after three frames; mixed128 -- three eager
    steps of the single 128^3 two-block simulation with the MIXED solve (single-simulation
    kernels, one symbol per multigrid level kind)."""
    from tests.scenes import FPS, dam_break, random_request, two_blocks
    a, b = pxl.DeviceBuffer(CALIB_BYTES), pxl.DeviceBuffer(CALIB_BYTES)
    b.copy_from(a)
    a.free()
    b.free()
    if args.probe_kind == "dam":
        sim = pxl.FluidState(dam_break(128, iters=80), FPS, seed=0, solver="cg", cg_iterations=80)
        sim.step(1)
        sim.profile_steps(2)
        sim.close()
        return {"probe": "dam"}
    if args.probe_kind == "mixed128":
        sim = pxl.FluidState(two_blocks(128), FPS, seed=0, solver="mixed", use_graph=False)
        sim.step(3)
        sim.close()
        return {"probe": "mixed128"}
    if args.probe_kind == "frame":
        f = dam_break(128, iters=80)
        sim = pxl.FluidState(f, FPS, seed=0, solver="cg", cg_iterations=80)
        sim.step(3 * f["substeps"])
        sim.profile_surface(1)
        sim.close()
        return {"probe": "frame"}
    sizes = tuple(int(v) for v in args.sizes.split(","))
    reqs = [random_request(i, sizes=sizes) for i in range(args.profile_batch)]
    with pxl.Batch(reqs, FPS, seeds=list(range(len(reqs)))) as bat:
        prof = bat.profile_steps(args.probe_steps)
    return {"probe": {k["kernel"]: k["alg_bytes"] / k["launches"] for k in prof}}


def bench_batch(args, pxl):
    """The fixed batch of the first --profile-batch stream requests: particle-steps/s of
    --steps batched steps after --warmup steps, and the per-kernel roofline."""
    from tests.scenes import random_request
    sizes = tuple(int(v) for v in args.sizes.split(","))
    reqs = [random_request(i, sizes=sizes) for i in range(args.profile_batch)]
    roofline, kernels, rate = batch_roofline(pxl, reqs, max(args.warmup, 1), args.steps,
                                             inner_iterations=args.inner_iterations)
    return {
        "metric": METRIC, "value": rate, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (random liquid boxes, seed = request index)",
        "config": {"workload": f"fixed batch of the first {len(reqs)} serving-stream requests, MIXED to 1e-6"},
        "roofline": roofline, "cpu_baseline": None, "kernels": kernels,
    }


def bench_tune(args, pxl):
    """The batch autotuner on the fixed batch of the first --profile-batch stream requests:
    coordinate descent over the batched launch families' persistent grids (each maximum
    halved and doubled, --steps timed steps per trial); value = the tuned batch's
    particle-steps/s.  --grids-out writes the chosen grids as a JSON cache that --grids loads."""
    from tests.scenes import random_request
    sizes = tuple(int(v) for v in args.sizes.split(","))
    reqs = [random_request(i, sizes=sizes) for i in range(args.profile_batch)]
    steps = min(args.steps, 10)
    r = pxl.autotune_batch(reqs, steps=steps, warmup=2, cache=args.grids_out)
    total = sum(int(pxl.check(f)["n_particles"]) for f in reqs)
    return {
        "metric": METRIC, "value": total / (r["ms_per_step"] / 1e3), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": steps, "warmup": 2, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (random liquid boxes or spheres, seed = request index)",
        "config": {"workload": f"batch autotuner on the first {len(reqs)} serving-stream requests, MIXED to 1e-6",
                   "grids": r["grids"], "shapes": r["shapes"], "baseline_ms_per_step": r["baseline_ms_per_step"],
                   "baseline_value": total / (r["baseline_ms_per_step"] / 1e3), "trials": r["trials"],
                   "cache": args.grids_out},
        "roofline": None, "cpu_baseline": None,
    }


def cpu_serve(args, reqs):
    """The C oracle in reference mode (the reference's red-black Gauss-Seidel, 80 sweeps) on the
    stream's first --cpu-requests requests, one step each, the simulations spread over
    --cpu-threads host threads (each simulation on one thread; the oracle releases the
    interpreter lock while it runs).  Oracle construction and seeding are not timed."""
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    from tests.scenes import oracle_sim
    sample = reqs[:args.cpu_requests]
    threads = max(1, min(args.cpu_threads, os.cpu_count() or 1))
    with ThreadPoolExecutor(threads) as ex:
        sims = list(ex.map(lambda f: oracle_sim(oracle, f), sample))
    sims.sort(key=lambda s: -s.P)  # largest first: the threads finish together
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda s: s.step_rbgs(80), sims))
    wall = time.perf_counter() - t0
    ps = sum(s.P for s in sims)
    return {"value": ps / wall, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"one reference-mode (RBGS-80) step of each of the first {len(sims)} stream requests "
                      f"({ps} particle-steps), C oracle, {threads} host threads, one simulation per thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="dam, mixed256, obstacles, batch: timed steps (default 60, obstacles 200); serve: not used "
                         "(each request is 60 steps)")
    ap.add_argument("--warmup", type=int, default=5, help="untimed warm-up steps")
    ap.add_argument("--workload", default="serve",
                    choices=["serve", "dam", "batch", "mixed256", "obstacles", "frame", "reference64", "agents", "tune",
                             "probe"])
    ap.add_argument("--grids", default=None, help="JSON cache of batched launch grids to load first (autotuner output)")
    ap.add_argument("--grids-out", default=None, help="tune: write the chosen grids to this JSON file")
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--sizes", default="32,48,64,96,128")
    ap.add_argument("--load", type=float, default=0.7,
                    help="serve: Poisson arrival rate of the latency run as a fraction of the measured capacity")
    ap.add_argument("--max-sims", type=int, default=512, help="serve: most simulations in the batch")
    ap.add_argument("--small-batch", type=int, default=8, help="serve: batch limit of the small-batch run (0: skip)")
    ap.add_argument("--small-requests", type=int, default=96, help="serve: requests of the small-batch run")
    ap.add_argument("--inner-iterations", type=int, default=9,
                    help="serve, batch: MG-CG iterations of the first refinement pass (later passes 2, 4, ...)")
    ap.add_argument("--budget", type=float, default=200e9)
    ap.add_argument("--plain-every", type=int, default=8, help="plain path on every k-th request")
    ap.add_argument("--profile-batch", type=int, default=64)
    ap.add_argument("--probe-steps", type=int, default=2, help="probe: event-timed batched steps")
    ap.add_argument("--probe-kind", default="serve", choices=["serve", "dam", "frame", "mixed128"], help="probe: which launches")
    ap.add_argument("--profile-every", type=int, default=40,
                    help="serve: every k-th batched step of the capacity run is event-timed per launch (roofline)")
    ap.add_argument("--traffic", type=int, default=1, help="serve: measure the dominant kernel's HBM traffic with "
                                                            "rocprofv3 counters in two child passes (0 = off)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="dam: CPU oracle sample length; serve: 0 skips the CPU baseline")
    ap.add_argument("--cpu-requests", type=int, default=192, help="serve: requests in the CPU oracle sample")
    ap.add_argument("--inner-max", type=int, default=32,
                    help="mixed256, obstacles: adaptive MG-CG pass length up to this many iterations (0 = fixed)")
    ap.add_argument("--cpu-threads", type=int, default=16, help="serve: host threads of the CPU oracle")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"obstacles": 200, "reference64": 100, "agents": 100}.get(args.workload, 60)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: every box here has one MI355X and this build has no "
              "multi-GPU path (requests shard whole simulations per GPU by design); run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import pxlengine_amd as pxl
    if args.grids:
        pxl.load_batch_grids(args.grids)
    run = {"tune": bench_tune, "serve": bench_serve, "dam": bench_dam, "batch": bench_batch, "mixed256": bench_mixed256,
           "obstacles": bench_obstacles, "frame": bench_frame, "reference64": bench_reference64,
           "agents": bench_agents, "probe": bench_probe}[args.workload]
    line = run(args, pxl)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
