"""C ABI on the CPU: the library loads, exports every declared entry point, and its host half
(strict fluid-block parsing, Scene::validate's fluid checks, Fluid::grid, seeding) behaves
like the reference -- none of this touches a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from tests.scenes import EXAMPLE_DAM_BREAK, dam_break

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def pxl():
    import pxlengine_amd
    pxlengine_amd.load_library()
    return pxlengine_amd


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "pxlflip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:pxlflip|pxlgrid)_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(pxl):
    lib = ctypes.CDLL(pxl.flip.LIB_PATH)
    names = _declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    # and the Python binding declares a signature for each of them
    assert set(names) == set(pxl.flip._SIGNATURES)


def _err(pxl, fluid, **kw):
    with pytest.raises(pxl.FluidError) as e:
        pxl.check(fluid, **kw)
    return e.value


def test_fluids_validate_reference_case(pxl):
    """crates/pxl-scene/src/lib.rs:886-899."""
    ok = {"name": "w", "bounds": [[0, 0, 0], [2, 1, 1]], "resolution": 32,
          "fill": [{"min": [0.1, 0.1, 0.1], "max": [0.5, 0.5, 0.5]}]}
    info = pxl.check(ok)
    assert info["n"] == (32, 16, 16)
    assert abs(info["h"] - 0.0625) < 1e-6
    outside = dict(ok, fill=[{"min": [0.1, 0.1, 0.1], "max": [0.5, 0.5, 1.5]}])
    e = _err(pxl, outside)
    assert "inside" in str(e) and e.code == pxl.flip.ERR_INVALID
    assert str(e) == "fluids[0] 'w': each fill block must have max > min and lie inside `bounds`"


def test_defaults_match_reference(pxl):
    """Defaults of crates/pxl-scene/src/lib.rs:96-119 (resolution 64, 8 ppc, 8 substeps,
    80 pressure iterations, gravity -9.81 y) and dt = f32(1/(fps*substeps))."""
    info = pxl.check({"name": "d", "bounds": [[0, 0, 0], [1, 1, 1]], "fill": [{"min": [0, 0, 0], "max": [0.5, 0.5, 0.5]}]})
    assert info["n"] == (64, 64, 64)
    assert info["substeps"] == 8 and info["pressure_iterations"] == 80
    assert info["gravity"] == (0.0, np.float32(-9.81), 0.0)
    assert info["flip"] == np.float32(0.95)
    assert info["dt"] == np.float32(1 / (24 * 8))
    nx, ny, nz = info["n"]
    assert info["n_faces"] == (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1)
    assert info["n_cells"] == nx * ny * nz


@pytest.mark.parametrize("change,message", [
    ({"resolution": 8}, "resolution 16..256, particles_per_cell 1..64, substeps 1..64, surface_refine 1..3, flip 0..1"),
    ({"resolution": 300}, "resolution 16..256"),
    ({"particles_per_cell": 0}, "particles_per_cell 1..64"),
    ({"substeps": 65}, "substeps 1..64"),
    ({"surface_refine": 4}, "surface_refine 1..3"),
    ({"flip": 1.5}, "flip 0..1"),
    ({"pressure_iterations": 0}, "flip 0..1"),
    ({"fill": []}, "needs at least one `fill` block of liquid"),
    ({"bounds": [[0, 0, 0], [1, 0, 1]]}, "bounds must be [min, max] with max > min on every axis"),
    ({"fill": [{"min": [0.5, 0, 0], "max": [0.4, 1, 1]}]}, "each fill block must have max > min"),
])
def test_validation_messages(pxl, change, message):
    f = dict(dam_break(32), **change)
    e = _err(pxl, f)
    assert message in str(e)
    assert str(e).startswith("fluids[0] 'dam32': ")


def test_grid_too_large(pxl):
    # 256 x 256 x 256 = 16.7M cells > 16,000,000 (lib.rs:731-734)
    f = dict(dam_break(256))
    assert "grid too large (lower resolution)" in str(_err(pxl, f))
    f["bounds"] = [[0, 0, 0], [1, 1, 0.9]]
    f["fill"] = [{"min": [0, 0, 0], "max": [0.4, 0.6, 0.9]}]
    assert pxl.check(f)["n"] == (256, 256, 230)


def test_strict_json(pxl):
    """deny_unknown_fields (lib.rs:50-51, 88): typos are rejected by name (cf. the reference
    test rejects_typos); required fields and types are enforced."""
    e = _err(pxl, dict(dam_break(32), resolutoin=32))
    assert "unknown field `resolutoin`" in str(e)
    e = _err(pxl, {"name": "x", "bounds": [[0, 0, 0], [1, 1, 1]], "fill": [{"min": [0, 0, 0], "max": [1, 1, 1], "vel": [0, 0, 0]}]})
    assert "unknown field `vel`" in str(e)
    assert "missing field `name`" in str(_err(pxl, {"bounds": [[0, 0, 0], [1, 1, 1]], "fill": []}))
    assert "missing field `fill`" in str(_err(pxl, {"name": "x", "bounds": [[0, 0, 0], [1, 1, 1]]}))
    assert "expected u32" in str(_err(pxl, dict(dam_break(32), resolution=32.5)))
    assert "expected u32" in str(_err(pxl, dict(dam_break(32), resolution=-1)))
    assert "invalid length 2" in str(_err(pxl, dict(dam_break(32), gravity=[0, 1])))
    assert "duplicate field" in str(_err(pxl, '{"name":"a","name":"b","bounds":[[0,0,0],[1,1,1]],"fill":[]}'))
    assert "invalid fluid" in str(_err(pxl, '{"name": "a", '))
    # the rendering block of examples/dam-break.json is accepted
    assert pxl.check(EXAMPLE_DAM_BREAK)["n_particles"] == 139200


def test_fps_and_empty_fill(pxl):
    assert str(_err(pxl, dam_break(32), fps=0.0)) == "render.fps must be > 0"
    # a fill block inside bounds that contains no cell centre (fluid.rs:149)
    f = dict(dam_break(32), fill=[{"min": [0.5, 0.5, 0.5], "max": [0.51, 0.51, 0.51]}])
    assert str(_err(pxl, f)) == "fluid 'dam32': fill blocks contain no grid cells"


@pytest.mark.parametrize("fluid,seed", [(dam_break(24, ppc=3), 0), (EXAMPLE_DAM_BREAK, 1)], ids=["dam24", "example"])
def test_host_seeding_matches_oracle(pxl, orc, fluid, seed):
    from tests.scenes import oracle_sim
    a = pxl.seed_particles(fluid, seed)
    sim = oracle_sim(orc, fluid, seed=seed)
    assert a.shape == (sim.P, 8)
    np.testing.assert_array_equal(a[:, 0:3], sim.x)
    np.testing.assert_array_equal(a[:, 4:7], sim.v)
    assert not a[:, 3].any() and not a[:, 7].any()


def test_multiple_blocks_with_velocity(pxl, orc):
    from tests.scenes import oracle_sim, two_blocks
    f = two_blocks(20)
    a = pxl.seed_particles(f, 4)
    sim = oracle_sim(orc, f, seed=4)
    np.testing.assert_array_equal(a[:, 0:3], sim.x)
    np.testing.assert_array_equal(a[:, 4:7], sim.v)
    assert set(np.unique(a[:, 4])) == {np.float32(1.5), np.float32(-1.5)}


def test_sphere_fill_seeding_matches_oracle(pxl, orc):
    """Sphere fills (an extension): strict parsing, bounds checks, seeding equal to the oracle."""
    from tests.scenes import oracle_sim
    f = dict(dam_break(24, ppc=3), fill=[{"center": [0.5, 0.4, 0.5], "radius": 0.25, "velocity": [0, -1, 0]},
                                        {"min": [0.05, 0.05, 0.05], "max": [0.3, 0.3, 0.3]}])
    a = pxl.seed_particles(f, 3)
    sim = oracle_sim(orc, f, seed=3)
    assert a.shape == (sim.P, 8) and sim.P > 0
    np.testing.assert_array_equal(a[:, 0:3], sim.x)
    np.testing.assert_array_equal(a[:, 4:7], sim.v)
    bad = dict(f, fill=[{"center": [0.5, 0.5, 0.5], "radius": 0.6}])
    assert "each fill sphere must have radius > 0 and lie inside `bounds`" in str(_err(pxl, bad))
    bad = dict(f, fill=[{"center": [0.5, 0.5, 0.5], "radius": 0.0}])
    assert "each fill sphere" in str(_err(pxl, bad))
    assert "missing field `radius`" in str(_err(pxl, dict(f, fill=[{"center": [0.5, 0.5, 0.5]}])))
    e = _err(pxl, dict(f, fill=[{"center": [0.5, 0.5, 0.5], "radius": 0.2, "min": [0, 0, 0]}]))
    assert "unknown field `min`" in str(e)


def test_obstacles_parse_validate_and_seed(pxl, orc):
    """Obstacles (an extension): strict parsing and checks; no liquid is seeded in obstacle
    cells (host seeding equal to the oracle's)."""
    from tests.scenes import obstacle_scene, obstacle_slab, oracle_sim
    f = obstacle_scene(20)
    a = pxl.seed_particles(f, 3)
    sim = oracle_sim(orc, f, seed=3)
    plain = oracle_sim(orc, dict(f, obstacles=[]), seed=3)
    assert a.shape == (sim.P, 8) and 0 < sim.P < plain.P
    np.testing.assert_array_equal(a[:, 0:3], sim.x)
    np.testing.assert_array_equal(a[:, 4:7], sim.v)
    slab = obstacle_slab(64)
    assert pxl.check(slab)["n_particles"] == oracle_sim(orc, slab).P
    msg = "each obstacle must have radius > 0 and its center inside `bounds`"
    assert msg in str(_err(pxl, dict(f, obstacles=[{"center": [0.5, 0.5, 0.5], "radius": 0}])))
    assert msg in str(_err(pxl, dict(f, obstacles=[{"center": [1.5, 0.5, 0.5], "radius": 0.1}])))
    assert "missing field `radius`" in str(_err(pxl, dict(f, obstacles=[{"center": [0.5, 0.5, 0.5]}])))
    e = _err(pxl, dict(f, obstacles=[{"center": [0.5, 0.5, 0.5], "radius": 0.1, "velocity": [0, 0, 0]}]))
    assert "unknown field `velocity`" in str(e)
    assert "expected a sequence" in str(_err(pxl, dict(f, obstacles={"center": [0.5, 0.5, 0.5]})))


def test_cfl_field(pxl):
    """CFL cap (an extension): `cfl` parses as a number, must be >= 0, 0 means off."""
    assert pxl.check(dict(dam_break(24), cfl=0.5))["n_particles"] == pxl.check(dam_break(24))["n_particles"]
    assert "cfl must be >= 0" in str(_err(pxl, dict(dam_break(24), cfl=-1.0)))
    assert "invalid type" in str(_err(pxl, dict(dam_break(24), cfl="fast")))


def test_device_bytes_reported(pxl):
    small = pxl.check(dam_break(32))["device_bytes"]
    mixed = pxl.check(dam_break(32), solver="mixed")["device_bytes"]
    assert 0 < small < mixed


def test_opts_validation(pxl):
    with pytest.raises(pxl.FluidError):
        pxl.check(dam_break(32), solver=7)
    with pytest.raises(pxl.FluidError):
        pxl.check(dam_break(32), solver="mixed", tolerance=0.0)


def test_server_options_are_validated_before_the_device(pxl):
    with pytest.raises(pxl.FluidError) as e:
        pxl.Server(max_sims=0)
    assert e.value.code == pxl.flip.ERR_INVALID and "max_sims" in str(e.value)
    with pytest.raises(pxl.FluidError) as e:
        pxl.Server(tolerance=0.0)
    assert e.value.code == pxl.flip.ERR_INVALID


def test_server_without_a_device_fails_cleanly(pxl):
    try:
        srv = pxl.Server()
    except pxl.FluidError as e:  # no GPU in this process
        assert e.code == pxl.flip.ERR_DEVICE and str(e)
    else:
        srv.close()


def test_batch_grid_settings(pxl, tmp_path):
    """The batched launch families' persistent-grid maxima (host-only settings): the measured
    defaults, validation (multiples of 8 in [8, 16384], known families) and the JSON cache the
    autotuner writes."""
    lib = pxl.load_library()
    saved = pxl.batch_grids()
    assert saved == {"band": 4096, "particle": 2048, "tile": 2048, "mg_setup": 2048, "mg_list": 512,
                     "mg_tail": 512, "solver": 1024}
    try:
        pxl.set_batch_grids({"tile": 1024, "solver": 8, "band": 16384})
        g = pxl.batch_grids()
        assert (g["tile"], g["solver"], g["band"], g["mg_list"]) == (1024, 8, 16384, 512)
        for bad in (0, 4, 1001, 16392):
            with pytest.raises(pxl.FluidError):
                pxl.set_batch_grids({"mg_list": bad})
        assert pxl.batch_grids()["mg_list"] == 512
        with pytest.raises(KeyError):
            pxl.set_batch_grids({"tiles": 8})
        out = ctypes.c_uint32()
        assert lib.pxlflip_batch_set_grid(len(pxl.GRID_FAMILIES), 8) == -1
        assert lib.pxlflip_batch_set_grid(-1, 8) == -1
        assert lib.pxlflip_batch_get_grid(len(pxl.GRID_FAMILIES), ctypes.byref(out)) == -1
        assert lib.pxlflip_batch_get_grid(0, None) == -1
        path = str(tmp_path / "grids.json")
        pxl.save_batch_grids(path, note="test")
        pxl.set_batch_grids(saved)
        assert pxl.load_batch_grids(path) == g
        assert pxl.batch_grids() == g
    finally:
        pxl.set_batch_grids(saved)
    assert pxl.batch_grids() == saved


def test_material_fields_are_strict(pxl):
    """A fluid's `materials` follow the reference's Material / MaterialInputs
    (crates/pxl-scene/src/lib.rs:365-444, deny_unknown_fields): known fields with their types are
    accepted, anything else is refused with serde's wording."""
    base = dict(dam_break(24))
    ok = [{"inputs": {"base_color": [0.1, 0.2, 0.3], "base_color_texture": "water.png", "specular_roughness": 0,
                      "normal_texture": None, "uv_scale": [2, 2], "transmission_weight": 1, "specular_ior": 1.33,
                      "transmission_color": [0.45, 0.85, 0.92], "transmission_depth": 0.8, "emission_luminance": 0,
                      "normal_scale": 1, "base_metalness": 0, "emission_color": [1, 1, 1]},
           "program": "wet", "role": "shadow_catcher"}, {}, {"role": "holdout", "program": None}]
    assert pxl.check(dict(base, materials=ok))["n_particles"] == pxl.check(base)["n_particles"]
    cases = [
        ([{"inputs": {"transmision_weight": 1}}], "unknown field `transmision_weight`, expected one of `base_color`"),
        ([{"colour": [1, 1, 1]}], "unknown field `colour`, expected one of `inputs`, `program`, `role`"),
        ([{"role": "glass"}], "unknown variant `glass`, expected one of `surface`, `shadow_catcher`, `holdout`"),
        ([{"role": 1}], "invalid type: integer `1`, expected enum Role"),
        ([{"inputs": {"base_color": [1, 1]}}], "invalid length 2, expected an array of length 3"),
        ([{"inputs": {"uv_scale": [1, 1, 1]}}], "invalid length 3, expected an array of length 2"),
        ([{"inputs": {"specular_ior": "1.33"}}], "expected f32"),
        ([{"inputs": {"normal_texture": 3}}], "expected a string"),
        ([{"program": 7}], "expected a string"),
        ([{"inputs": []}], "expected struct MaterialInputs"),
        (["water"], "expected struct Material"),
    ]
    for mats, msg in cases:
        e = _err(pxl, dict(base, materials=mats))
        assert msg in str(e), (mats, str(e))
        assert str(e).startswith("invalid fluid: "), str(e)


def test_batch_rejects_members_beyond_the_work_list_field(pxl):
    """A batch work-list entry keeps a member's workgroup in 20 bits, so a member needing 2^20
    or more workgroups of 256 particles (over 268,435,456 particles: a full 250^3 grid at 20
    particles per cell, inside the reference's 16M-cell limit) is refused before any device
    work, instead of overflowing into the member bits."""
    big = {"name": "big", "bounds": [[0, 0, 0], [1, 1, 1]], "resolution": 250,
           "fill": [{"min": [0, 0, 0], "max": [1, 1, 1]}], "particles_per_cell": 20}
    info = pxl.check(big)
    assert info["n_particles"] == 248 ** 3 * 20 > 1 << 28
    with pytest.raises(pxl.FluidError) as e:
        pxl.Batch([big])
    assert e.value.code == pxl.flip.ERR_INVALID
    assert "too large for batched execution" in str(e.value) and "1048576" in str(e.value)
    ok = dict(big, particles_per_cell=16)  # 244M particles: 953,312 workgroups fit
    assert pxl.check(ok)["n_particles"] < 1 << 28


def test_batch_shape_settings(pxl, tmp_path):
    """The workgroup sizes of the families compiled in two sizes (host-only settings): the
    defaults, validation (only the compiled sizes) and the JSON cache round trip."""
    saved = pxl.batch_shapes()
    assert saved == {"tile_lanes": 512, "band_lanes": 256}
    assert pxl.SHAPE_FAMILIES == {"tile_lanes": (512, 1024), "band_lanes": (256, 512)}
    lib = pxl.load_library()
    try:
        pxl.set_batch_shapes({"tile_lanes": 512, "band_lanes": 512})
        assert pxl.batch_shapes() == {"tile_lanes": 512, "band_lanes": 512}
        for bad in (256, 768, 0):
            with pytest.raises(pxl.FluidError):
                pxl.set_batch_shapes({"tile_lanes": bad})
        with pytest.raises(KeyError):
            pxl.set_batch_shapes({"tiles": 512})
        assert lib.pxlflip_batch_set_shape(2, 512) == -1 and lib.pxlflip_batch_set_shape(-1, 512) == -1
        path = str(tmp_path / "tuned.json")
        pxl.save_batch_grids(path)
        pxl.set_batch_shapes(saved)
        pxl.load_batch_grids(path)
        assert pxl.batch_shapes() == {"tile_lanes": 512, "band_lanes": 512}
    finally:
        pxl.set_batch_shapes(saved)
    assert pxl.batch_shapes() == saved
