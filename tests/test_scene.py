"""Whole scene documents on the CPU: Scene::from_json / Scene::validate for what the fluid path
reads (crates/pxl-scene/src/lib.rs:646-745) and the `pxl check` summary line
(crates/pxl-cli/src/check.rs:98-104), through the C ABI."""
import json
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
DAM = os.path.join(HERE, "golden", "scenes", "dam-break.json")  # examples/dam-break.json of the reference


@pytest.fixture(scope="module")
def pxl():
    import pxlengine_amd
    return pxlengine_amd


def _err(pxl, scene, **kw):
    with pytest.raises(pxl.FluidError) as e:
        pxl.check_scene(scene, **kw)
    assert e.value.code == pxl.flip.ERR_INVALID
    return str(e.value)


def test_reference_test_fluids_validate(pxl):
    """The reference's own test fluids_validate (lib.rs:886-899), as whole scenes."""
    ok = ('{"version":1,"nodes":[{"type":"camera"}],"fluids":[{"name":"w",'
          '"bounds":[[0,0,0],[2,1,1]],"resolution":32,"fill":[{"min":[0.1,0.1,0.1],"max":[0.5,0.5,0.5]}]}]}')
    s = pxl.check_scene(ok)
    assert s["fluids"] == 1 and s["line"] == "ok: 1 nodes, frames 0..0, 0 simulation(s), 1 fluid(s)"
    fluid = json.loads(ok)["fluids"][0]
    info = pxl.check(fluid)
    assert info["n"] == (32, 16, 16) and abs(info["h"] - 0.0625) < 1e-6
    outside = ok.replace('"max":[0.5,0.5,0.5]', '"max":[0.5,0.5,1.5]')
    assert "inside" in _err(pxl, outside)
    glass = ('{"version":1,"nodes":[{"type":"camera"},{"type":"mesh","geometry":{"std":"sphere"},'
             '"materials":[{"inputs":{"transmission_weight":1,"specular_ior":1.33}}]}]}')
    assert pxl.check_scene(glass)["fluids"] == 0


def test_example_dam_break_whole_scene(pxl):
    """examples/dam-break.json loaded whole: 7 nodes, the animation's frames 0..119, one fluid
    whose block seeds the published 139,200 particles (docs/ROADMAP.md:83) at render.fps 24
    and render.seed 0 (the defaults: the scene sets neither)."""
    text = open(DAM).read()
    s = pxl.check_scene(text)
    assert s["line"] == "ok: 7 nodes, frames 0..119, 0 simulation(s), 1 fluid(s)"
    assert (s["version"], s["nodes"], s["fluids"], s["fps"], s["seed"], s["frames"]) == (1, 7, 1, 24.0, 0, (0, 119))
    assert pxl.check(json.loads(text)["fluids"][0], s["fps"], seed=s["seed"])["n_particles"] == 139200


def _scene(**top):
    base = json.loads(open(DAM).read())
    base.update(top)
    return json.dumps(base)


@pytest.mark.parametrize("change,message", [
    (lambda s: s.update(colour=1), "invalid scene: unknown field `colour`, expected one of `version`, `nodes`"),
    (lambda s: s.pop("version"), "invalid scene: missing field `version`"),
    (lambda s: s.update(version=2), "scene version 2 is not supported (expected 1)"),
    (lambda s: s.update(nodes=[n for n in s["nodes"] if n["type"] != "camera"]), "scene has no camera node"),
    (lambda s: s["nodes"].append({"type": "spotlight"}), "unknown variant `spotlight`"),
    (lambda s: s["render"].update(fps=0), "render.fps must be > 0"),
    (lambda s: s["render"].update(fsp=30), "invalid scene: unknown field `fsp`, expected one of `resolution`"),
    (lambda s: s["render"].update(resolution=[0, 720]), "render.resolution must be within 1..16384, got [0,720]"),
    (lambda s: s["render"].update(view="filmic"), "unknown variant `filmic`, expected one of `agx`, `aces`, `standard`"),
    (lambda s: s["render"].update(seed=-1), "expected u32"),
    (lambda s: s["render"].update(motion_blur=1.5), "render.motion_blur is the shutter fraction of a frame, 0..1"),
    (lambda s: s.pop("programs"), "a material uses program 'tiles' but the scene has no `programs` block"),
    (lambda s: s["nodes"][0]["materials"][0].update(program="9tiles"), "material program '9tiles' must be the name"),
    (lambda s: s["fluids"][0].update(resolutoin=64), "invalid scene: unknown field `resolutoin`"),
    (lambda s: s["fluids"][0]["materials"][0]["inputs"].update(ior=1.3), "invalid scene: unknown field `ior`"),
    (lambda s: s["fluids"].append(dict(s["fluids"][0], name="second", flip=2.0)),
     "fluids[1] 'second': resolution 16..256"),
    (lambda s: s["fluids"][0].update(resolution=256, bounds=[[0, 0, 0], [1, 1, 1]],
                                     fill=[{"min": [0, 0, 0], "max": [0.5, 0.5, 0.5]}]),
     "fluids[0] 'water': grid too large (lower resolution)"),
])
def test_scene_errors(pxl, change, message):
    s = json.loads(open(DAM).read())
    change(s)
    assert message in _err(pxl, json.dumps(s))


def test_render_block_and_animation_range(pxl):
    """render.fps / render.seed reach the fluid path; the frame range follows Animation::end
    (frame_end, else the last key or orbit frame, at least frame_start); max_cells raises the
    16M-cell limit for every fluid of the scene (an extension)."""
    s = json.loads(open(DAM).read())
    s["render"].update(fps=30, seed=7)
    s["animation"] = {"frame_start": 5, "tracks": [{"node": "cam", "path": "translation", "keys": [[0, [0, 0, 0]], [40.5, [1, 0, 0]]]},
                                                   {"node": "cam", "path": "target", "orbit": {"center": [0, 0, 0], "frames": [0, 60]}}]}
    out = pxl.check_scene(json.dumps(s))
    assert (out["fps"], out["seed"], out["frames"]) == (30.0, 7, (5, 60))
    s["animation"] = {"frame_start": 3}
    assert pxl.check_scene(json.dumps(s))["frames"] == (3, 3)
    big = json.loads(open(DAM).read())
    big["fluids"][0].update(resolution=256, bounds=[[0, 0, 0], [1, 1, 1]], fill=[{"min": [0, 0, 0], "max": [0.5, 0.5, 0.5]}])
    assert pxl.check_scene(json.dumps(big), max_cells=1 << 25)["fluids"] == 1


def test_every_reference_example_scene_loads(pxl):
    """The scene-level checks accept every example scene of the reference that is here (this
    container only: the reference is not on the GPU box)."""
    import glob
    paths = glob.glob("/root/reference/examples/*.json") + glob.glob("/root/reference/examples/vfx/*.json")
    if not paths:
        pytest.skip("the reference's examples are not present")
    for p in paths:
        assert pxl.check_scene(open(p).read())["line"].startswith("ok: "), p
