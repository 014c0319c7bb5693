"""Writes the oracle's reference-mode trajectories of the dam-break configurations as digests.

The oracle (oracle/flip_oracle.c, the C restatement of shaders/fluid.slang) runs the dam break
with the reference's red-black Gauss-Seidel pressure (80 sweeps) and the CFL cap (1 cell) for
100 steps.  After each recorded step it writes the BLAKE2b-128 digest of:
  x, v       particle positions and velocities, float32 [P, 3] in seeding order
  pressure   float32 [C]
  vel        face velocities, float32 [F] (u, v, w blocks)
  count      the step's per-cell particle count (its last substep's p2g), uint32 [C]
  bins       cell of every particle, counts, offsets and cell-grouped ids of the particles'
             positions after the step (the counting sort), uint32, concatenated
and the step's CFL substep count.  A trajectory of one of these sizes is too large to commit
(14.7 MB of particles per step at 64^3), so the GPU test hashes its own state the same way and
compares digests, step by step.

    python tests/golden/make_trajectories.py dam64     # ~2 min
    python tests/golden/make_trajectories.py dam128    # ~30 min
    python tests/golden/make_trajectories.py dam128 --checkpoint /tmp/dam128.npz   # resumable

With --checkpoint the oracle's carried state (particles and the warm-start pressure: every other
array is rebuilt by the next step) and the digests so far are saved after every recorded step,
and a later run with the same path continues from them, bitwise as an uninterrupted run.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from tests.scenes import FPS, dam_break, oracle_sim  # noqa: E402

ITERS = 80
CFL = 1.0
STEPS = 100
SEED = 0

# configuration -> (grid resolution, recorded steps)
CONFIGS = {
    "dam64": (64, list(range(1, STEPS + 1))),
    "dam128": (128, list(range(1, 11)) + list(range(20, STEPS + 1, 10))),
}


def scene(res: int) -> dict:
    return dict(dam_break(res, iters=ITERS), cfl=CFL)


def digest(*arrays) -> str:
    h = hashlib.blake2b(digest_size=16)
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def state_digests(x, v, pressure, vel, count, bins) -> dict:
    return {"x": digest(np.asarray(x, np.float32)), "v": digest(np.asarray(v, np.float32)),
            "pressure": digest(np.asarray(pressure, np.float32)), "vel": digest(np.asarray(vel, np.float32)),
            "count": digest(np.asarray(count, np.uint32)),
            "bins": digest(*[np.asarray(b, np.uint32) for b in bins])}


def fixture_path(name: str) -> str:
    return os.path.join(HERE, f"trajectory_{name}.json")


def oracle_steps(name: str, upto: int, checkpoint: str = None, done: dict = None):
    """Yields (step, substeps, digests) of the oracle trajectory of `name` up to `upto`; with a
    checkpoint, starts after the state saved there and saves the state after each recorded
    step (with `done`, the digests so far)."""
    import oracle
    res, record = CONFIGS[name]
    sim = oracle_sim(oracle, scene(res), seed=SEED)
    start = 1
    if checkpoint and os.path.exists(checkpoint):
        with np.load(checkpoint) as z:
            sim.set_particles(z["x"], z["v"])
            sim.pressure[:] = z["pressure"]
            start = int(z["step"]) + 1
    for s in range(start, upto + 1):
        k = sim.step_rbgs_capped(ITERS, CFL)
        if s in record:
            yield s, k, state_digests(sim.x, sim.v, sim.pressure, sim.vel, sim.count, sim.bin())
            if checkpoint:
                tmp = checkpoint + ".tmp.npz"
                np.savez(tmp, x=sim.x, v=sim.v, pressure=sim.pressure, step=s,
                         done=json.dumps(done if done is not None else {}))
                os.replace(tmp, checkpoint)


def main(name: str, checkpoint: str = None):
    res, record = CONFIGS[name]
    out = {"config": name, "scene": scene(res), "fps": FPS, "seed": SEED, "pressure_iterations": ITERS,
           "cfl": CFL, "solver": "rbgs", "digest": "blake2b-128", "steps": {}}
    if checkpoint and os.path.exists(checkpoint):
        with np.load(checkpoint) as z:
            out["steps"] = json.loads(str(z["done"]))
    t0 = time.time()
    for s, k, d in oracle_steps(name, max(record), checkpoint, out["steps"]):
        out["steps"][str(s)] = dict(d, substeps=k)
        print(f"{name} step {s}: {k} substeps, {time.time() - t0:.0f} s", flush=True)
    with open(fixture_path(name), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    ck = sys.argv[sys.argv.index("--checkpoint") + 1] if "--checkpoint" in sys.argv else None
    main(sys.argv[1], ck)
