"""Fluid blocks used by the tests and the bench, and an oracle simulation built from one.

The oracle side derives everything itself (grid, dt, seeding) from the block, so a product
mistake in parsing or seeding shows up as a mismatch instead of being copied over.
"""
from __future__ import annotations

import numpy as np

FPS = 24.0


def dam_break(res: int, substeps: int = 5, flip: float = 0.95, ppc: int = 8, iters: int = 80,
              fill_max=(0.4, 0.6, 1.0)) -> dict:
    """The dam break of the configs: unit cube, liquid block x < 0.4, y < 0.6, at rest.
    substeps 5 at 24 fps gives dt = f32(1/120)."""
    return {
        "name": f"dam{res}", "bounds": [[0, 0, 0], [1, 1, 1]], "resolution": res,
        "fill": [{"min": [0, 0, 0], "max": list(fill_max)}],
        "particles_per_cell": ppc, "substeps": substeps, "flip": flip,
        "gravity": [0, -9.81, 0], "pressure_iterations": iters,
    }


# examples/dam-break.json's fluid block (the reference's only fluid scene).
EXAMPLE_DAM_BREAK = {
    "name": "water", "bounds": [[-1, 0, -0.5], [1, 1.2, 0.5]], "resolution": 64,
    "fill": [{"min": [-0.97, 0.03, -0.47], "max": [-0.35, 0.95, 0.47]}],
    "substeps": 8, "flip": 0.95,
    "materials": [{"inputs": {"transmission_weight": 1, "specular_ior": 1.33, "specular_roughness": 0,
                              "transmission_color": [0.45, 0.85, 0.92], "transmission_depth": 0.8}}],
}


def two_blocks(res: int, ppc: int = 8, iters: int = 40) -> dict:
    """Two liquid blocks moving toward each other (mixed-precision config shape)."""
    return {
        "name": f"collide{res}", "bounds": [[0, 0, 0], [1, 1, 1]], "resolution": res,
        "fill": [{"min": [0.02, 0.02, 0.02], "max": [0.35, 0.7, 0.98], "velocity": [1.5, 0, 0]},
                 {"min": [0.65, 0.02, 0.02], "max": [0.98, 0.7, 0.98], "velocity": [-1.5, 0, 0]}],
        "particles_per_cell": ppc, "substeps": 5, "flip": 0.95, "pressure_iterations": iters,
    }


def sphere_drop(res: int, ppc: int = 6, iters: int = 40) -> dict:
    """A liquid sphere (an extension) thrown downward onto a shallow pool."""
    return {
        "name": f"sphere{res}", "bounds": [[0, 0, 0], [1, 1, 1]], "resolution": res,
        "fill": [{"center": [0.45, 0.62, 0.5], "radius": 0.25, "velocity": [0.5, -2.0, 0.2]},
                 {"min": [0.02, 0.02, 0.02], "max": [0.98, 0.15, 0.98]}],
        "particles_per_cell": ppc, "substeps": 5, "flip": 0.95, "pressure_iterations": iters,
    }


def obstacle_scene(res: int, ppc: int = 4, iters: int = 40) -> dict:
    """Liquid around and above solid spheres (obstacles, an extension), thrown down onto them;
    the last obstacle overlaps the first."""
    return {
        "name": f"obstacles{res}", "bounds": [[0, 0, 0], [1, 1, 1]], "resolution": res,
        "fill": [{"min": [0.05, 0.05, 0.05], "max": [0.95, 0.6, 0.95], "velocity": [0.3, -1.5, 0.2]}],
        "obstacles": [{"center": [0.5, 0.3, 0.5], "radius": 0.15}, {"center": [0.25, 0.25, 0.7], "radius": 0.1},
                      {"center": [0.72, 0.4, 0.28], "radius": 0.12}, {"center": [0.6, 0.35, 0.55], "radius": 0.1}],
        "particles_per_cell": ppc, "substeps": 5, "flip": 0.95, "pressure_iterations": iters,
    }


def obstacle_slab(res: int = 192, seed: int = 5, solid_fraction: float = 0.05, ppc: int = 8) -> dict:
    """Obstacle-heavy scene on [0,2] x [0,1] x [0,2] (res cells along x and z, res/2 along y):
    solid spheres of radius 2-8 cells at seeded random centres until their summed volume
    reaches `solid_fraction` of the domain, and a liquid slab over the full width from y = 0.25
    to the top, falling onto them."""
    rng = np.random.default_rng(seed)
    h = 2.0 / res
    obst, vol = [], 0.0
    while vol < solid_fraction * 2.0 * 1.0 * 2.0:
        r = float(rng.uniform(2.0, 8.0)) * h
        c = [float(rng.uniform(r + h, 2.0 - r - h)), float(rng.uniform(r + h, 1.0 - r - h)),
             float(rng.uniform(r + h, 2.0 - r - h))]
        obst.append({"center": c, "radius": r})
        vol += 4.0 / 3.0 * np.pi * r ** 3
    return {
        "name": f"slab{res}", "bounds": [[0, 0, 0], [2, 1, 2]], "resolution": res,
        "fill": [{"min": [0.0, 0.25, 0.0], "max": [2.0, 1.0, 2.0]}], "obstacles": obst,
        "particles_per_cell": ppc, "substeps": 5, "flip": 0.95,
    }


def oracle_sim(orc, fluid: dict, fps: float = FPS, seed: int = 0):
    """Builds the oracle simulation for a fluid block (defaults as crates/pxl-scene/src/lib.rs:96-119)."""
    lo, hi = fluid["bounds"]
    n, h = orc.fluid_grid(lo, hi, fluid.get("resolution", 64))
    blocks = [("sphere", b["center"], b["radius"], b.get("velocity", [0, 0, 0])) if "center" in b
              else (b["min"], b["max"], b.get("velocity", [0, 0, 0])) for b in fluid["fill"]]
    obst = [(o["center"], o["radius"]) for o in fluid.get("obstacles", [])]
    x, v = orc.seed_particles(n, h, lo, blocks, fluid.get("particles_per_cell", 8), seed, obstacles=obst)
    dt = orc.dt_for(fps, fluid.get("substeps", 8))
    return orc.Sim(n, h, lo, dt, fluid.get("flip", 0.95), fluid.get("gravity", [0, -9.81, 0]), x, v,
                   obstacles=obst)


def aos8(x: np.ndarray, v: np.ndarray) -> np.ndarray:
    out = np.zeros((len(x), 8), np.float32)
    out[:, 0:3] = x
    out[:, 4:7] = v
    return out


def random_request(index: int, sizes=(32, 48, 64, 96, 128), ppc_range=(4, 8), spheres=True) -> dict:
    """A serving request: cubic grid n^3 with n drawn from `sizes`, one liquid region at a
    random place and size inside the interior -- a box, or with `spheres` half the time a
    sphere (an extension) -- ppc drawn from `ppc_range`, a random initial velocity.
    Deterministic in `index`."""
    rng = np.random.default_rng(1000 + index)
    n = int(rng.choice(sizes))
    sphere = spheres and bool(rng.integers(2))
    if sphere:
        r = rng.uniform(0.12, 0.3)
        c = rng.uniform(0.05 + r, 0.95 - r, 3)
        vel = rng.uniform(-1.0, 1.0, 3)
        fill = {"center": [float(a) for a in c], "radius": float(r), "velocity": [float(a) for a in vel]}
    else:
        lo = rng.uniform(0.05, 0.45, 3)
        ext = rng.uniform(0.2, 0.5, 3)
        hi = np.minimum(lo + ext, 0.95)
        vel = rng.uniform(-1.0, 1.0, 3)
        fill = {"min": [float(a) for a in lo], "max": [float(a) for a in hi], "velocity": [float(a) for a in vel]}
    return {
        "name": f"req{index}", "bounds": [[0, 0, 0], [1, 1, 1]], "resolution": n, "fill": [fill],
        "particles_per_cell": int(rng.integers(ppc_range[0], ppc_range[1] + 1)), "substeps": 5, "flip": 0.95,
    }
