#!/usr/bin/env python3
"""Throughput of the FLIP liquid step on one MI355X: particle-steps/s.

Workload: the single-instance dam break of the baseline configs -- unit cube at 128^3
(h = 1/128), liquid block x < 0.4, y < 0.6 seeded at 8 particles per cell (3,830,400
particles, fp32, at rest, seed 0), gravity -9.81, dt = f32(1/120), PIC/FLIP 0.95, and the
plain FP32 pressure CG with a fixed 80 iterations.  A step is one whole FLIP step of that
simulation (bin, P2G, grid ops, pressure, project, G2P) with its state resident in HBM.

Prints one JSON line.  `roofline` is for the kernel with the largest share of step time,
from HIP events bracketing each launch on the simulation's stream; `cpu_baseline` times
the CPU oracle (single thread) on a bounded sample of the same workload.
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


def algorithmic_bytes(name: str, P: int, F: int, C: int, Cf: int) -> float:
    """Minimal bytes one launch of a kernel must move: every logical array once, at its
    element width (fp32/int32 4 B, cell type 1 B, stencil 2 B, particle x+v 24 B)."""
    table = {
        "clear_grid": 8 * F + 4 * C,
        "bin_count": 12 * P + 8 * P + 4 * C,
        "scan_offsets": 8 * C,
        "scatter": 36 * P + 28 * P + 4 * C,
        "p2g": 24 * P + 8 * F,
        "faces": 16 * F,
        "cells": 4 * C + 4 * F + 7 * C + 4 * (C - Cf),
        "rbgs": C + 8 * Cf,
        "cg_start": 2 * C + 12 * Cf + 2 * C + 16 * Cf,
        "cg_iter": 2 * C + 40 * Cf,
        "project": 8 * F + 5 * C,
        "g2p": 28 * P + 8 * F + 28 * P,
    }
    return float(table.get(name, 0))


def measured_traffic(kernel: str, args):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary of this workload (profiles/), or None when there is none for it."""
    path = os.path.join(ROOT, "profiles", f"traffic_dam{args.res}_{args.solver}{args.iters}.json")
    if not os.path.exists(path):
        return None
    per = json.load(open(path))["per_dispatch"]
    for key in (f"k_{kernel}", f"k_{kernel}<float>", f"k_{kernel}_tile"):
        if key in per:
            return per[key]["hbm_bytes"]
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--res", type=int, default=128)
    ap.add_argument("--solver", default="cg", choices=["cg", "rbgs", "mixed"])
    ap.add_argument("--iters", type=int, default=80)
    ap.add_argument("--cpu-steps", type=int, default=2, help="oracle steps timed for cpu_baseline (0 = skip)")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: every box here has one MI355X and this build has no "
              "multi-GPU path (requests shard whole simulations per GPU by design); run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import pxlengine_amd as pxl
    from tests.scenes import FPS, dam_break

    fluid = dam_break(args.res, iters=args.iters)
    sim = pxl.FluidState(fluid, FPS, seed=0, solver=args.solver, cg_iterations=args.iters)
    P = sim.particle_count()
    info = sim.info
    for _ in range(args.warmup):
        sim.step(1)
    tuned = sim.autotune(2) if not args.no_tune else None  # launch shape from the autotuner
    t0 = time.perf_counter()
    sim.step(args.steps)  # returns when the GPU work is complete
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = P * args.steps / elapsed

    # per-kernel times from HIP events on the simulation's stream (extra, untimed steps)
    prof = sim.profile_steps(2)
    Cf = int((sim.read_field("type") == 1).sum())
    F, Cc = info["n_faces"], info["n_cells"]
    per_kernel = []
    for name, launches, total_ms in prof:
        avg_s = total_ms / 1e3 / launches
        b = algorithmic_bytes(name, P, F, Cc, Cf)
        per_kernel.append({"kernel": name, "launches_per_step": launches // 2, "avg_us": avg_s * 1e6,
                           "alg_bytes": b, "achieved_GBs": (b / avg_s / 1e9) if b else None})
    dom = per_kernel[0]
    achieved = dom["achieved_GBs"] or 0.0
    roofline = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS, "traffic": measured_traffic(dom["kernel"], args),
                "kernel": dom["kernel"], "avg_us": dom["avg_us"], "alg_bytes_per_launch": dom["alg_bytes"]}
    stats = sim.last_solve()
    sim.close()

    cpu = None
    if args.cpu_steps > 0:
        import oracle
        from tests.scenes import oracle_sim
        osim = oracle_sim(oracle, fluid)
        c0 = time.perf_counter()
        for _ in range(args.cpu_steps):
            if args.solver == "rbgs":
                osim.step_rbgs(args.iters)
            else:
                osim.step_cg(args.iters)
        c1 = time.perf_counter()
        cpu = {"value": osim.P * args.cpu_steps / (c1 - c0), "unit": "particle-steps/s", "cores": 1,
               "kind": "port",
               "sample": f"first {args.cpu_steps} steps of the same {args.res}^3 dam break ({osim.P} particles), "
                         f"{'RBGS' if args.solver == 'rbgs' else 'FP32 CG'} {args.iters} iterations, "
                         "C oracle, one thread"}
        del osim

    solver_desc = {"cg": f"plain FP32 Jacobi-PCG, {args.iters} fixed iterations",
                   "rbgs": f"reference red-black Gauss-Seidel, {args.iters} iterations",
                   "mixed": "FP32 CG inside FP64 refinement to 1e-6"}[args.solver]
    line = {
        "metric": "particle-steps/s batched vs plain at ≤1e-6 rel residual; p50/p99 sim latency",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded dam break, reference LCG seeding)",
        "config": {"workload": f"dam break {args.res}^3, single instance, {solver_desc}",
                   "grid": list(info["n"]), "particles": P, "dt": float(info["dt"]), "flip": float(info["flip"]),
                   "path": "plain (one simulation, one stream, hipGraph per step)",
                   "autotune": tuned},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": per_kernel,
        "fluid_cells": Cf,
        "last_solve": stats,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
