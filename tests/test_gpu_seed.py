"""Device seeding: the initial particles a simulation is created with (seeded by HIP kernels, the
LCG jumped ahead to each cell's first draw) equal the host restatement of the reference's seeding
loop (crates/pxl-render/src/fluid.rs:123-148) and the oracle's, bit for bit -- boxes, spheres,
several and overlapping blocks (the stream continues across blocks; overlaps seed twice),
obstacles, blocks touching the walls, and the members of a batch."""
import numpy as np
import pytest

from tests.scenes import (FPS, EXAMPLE_DAM_BREAK, dam_break, obstacle_scene, obstacle_slab, oracle_sim, random_request,
                          sphere_drop, two_blocks)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pxl():
    import pxlengine_amd
    return pxlengine_amd


OVERLAP = {"name": "overlap", "bounds": [[0, 0, 0], [1, 1, 1]], "resolution": 40, "particles_per_cell": 3,
           "fill": [{"min": [0.1, 0.1, 0.1], "max": [0.6, 0.5, 0.7], "velocity": [1, 0, 0]},
                    {"center": [0.5, 0.5, 0.5], "radius": 0.3, "velocity": [0, -1, 0]},
                    {"min": [0.0, 0.0, 0.0], "max": [1.0, 0.2, 1.0]}]}
SCENES = [(dam_break(32), 0), (EXAMPLE_DAM_BREAK, 1), (two_blocks(24), 7), (sphere_drop(28), 3),
          (obstacle_scene(24), 2), (OVERLAP, 9), (random_request(17), 17), (obstacle_slab(64), 0),
          (two_blocks(128), 5)]


@pytest.mark.parametrize("fluid,seed", SCENES, ids=[f["name"] for f, _ in SCENES])
def test_device_seeding_bitwise(pxl, orc, fluid, seed):
    print(f"seeding {fluid['name']} seed {seed}", flush=True)
    gpu = pxl.FluidState(fluid, FPS, seed=seed)
    got = gpu.read_particles()
    host = pxl.seed_particles(fluid, seed)
    assert got.shape == host.shape
    np.testing.assert_array_equal(got.view(np.uint32), host.view(np.uint32))
    sim = oracle_sim(orc, fluid, seed=seed)
    np.testing.assert_array_equal(got[:, 0:3], sim.x)
    np.testing.assert_array_equal(got[:, 4:7], sim.v)
    gpu.close()


def test_batch_members_seeded_bitwise(pxl):
    fluids = [random_request(i) for i in (0, 1, 2, 13)]
    with pxl.Batch(fluids, FPS, seeds=[4, 5, 6, 7]) as b:
        for k, (f, s) in enumerate(zip(fluids, [4, 5, 6, 7])):
            np.testing.assert_array_equal(b.read_particles(k).view(np.uint32),
                                          pxl.seed_particles(f, s).view(np.uint32))


def test_create_from_scene_uses_render_block(pxl):
    """fluids[i] of a whole scene (examples/dam-break.json, render.seed and render.fps set) is
    created with the scene's seed and fps: its particles are the fluid block's seeding for
    render.seed and its dt is f32(1 / (fps * substeps))."""
    import json
    import os
    s = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes", "dam-break.json")))
    s["render"].update(seed=7, fps=30)
    gpu = pxl.FluidState.from_scene(json.dumps(s), 0)
    fluid = s["fluids"][0]
    np.testing.assert_array_equal(gpu.read_particles().view(np.uint32), pxl.seed_particles(fluid, 7).view(np.uint32))
    assert gpu.info["dt"] == np.float32(1.0 / (30 * 8)) and gpu.fps == 30.0 and gpu.seed == 7
    gpu.step(1)
    gpu.close()
    with pytest.raises(pxl.FluidError, match="no fluids\\[1\\]"):
        pxl.FluidState.from_scene(json.dumps(s), 1)
