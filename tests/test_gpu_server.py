"""Batched serving on the GPU: every simulation's result in a batch equals its result alone
(bitwise: same kernels, same workgroup partition, same reduction order, its own convergence),
each pressure solve reaches the FP64 tolerance, admission respects the HBM budget, requests
arriving while others run join the batch, and cancellation retires a running request."""
import time

import numpy as np
import pytest

from tests.scenes import FPS, dam_break, obstacle_scene, random_request, two_blocks

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pxl():
    import pxlengine_amd
    return pxlengine_amd

WG = 32  # solver workgroups per simulation, fixed so the single-sim path can use the same partition


def _server(pxl, **kw):
    kw.setdefault("solver_workgroups", WG)
    return pxl.Server(**kw)


def _run(pxl, reqs, **kw):
    """reqs: [(fluid, steps)] submitted together; returns ([particles], [info], stats)."""
    with _server(pxl, **kw) as srv:
        ids = [srv.submit(f, FPS, seed=0, steps=n, keep_particles=True) for f, n in reqs]
        infos = [srv.wait(i, timeout=120) for i in ids]
        parts = [srv.read_particles(i) for i in ids]
        return parts, infos, srv.stats()


def test_batched_equals_single_sim_path(pxl):
    f = dam_break(24)
    (p_batch,), (info,), _ = _run(pxl, [(f, 3)])
    assert info["status"] == "done" and info["rel_residual"] <= 1e-6
    sim = pxl.FluidState(f, FPS, seed=0, solver="mixed", solver_workgroups=WG)
    sim.step(3)
    assert np.array_equal(p_batch, sim.read_particles())
    sim.close()


def test_alone_equals_in_the_mix(pxl):
    a, b, c = dam_break(32), two_blocks(24), random_request(3, sizes=(16, 40))
    d = random_request(1, sizes=(16, 40))  # a sphere fill
    e = obstacle_scene(24)
    assert "center" in d["fill"][0] and "min" in c["fill"][0]
    (pa,), _, _ = _run(pxl, [(a, 4)])
    parts, infos, st = _run(pxl, [(b, 2), (a, 4), (c, 5), (d, 3), (e, 4)])
    for info in infos:
        assert info["status"] == "done"
        assert 0 <= info["rel_residual"] <= 1e-6
    assert np.array_equal(pa, parts[1])
    assert st["completed"] == 5 and st["mean_batch"] > 1.0
    for k, (f, n) in ((2, (c, 5)), (3, (d, 3)), (4, (e, 4))):
        (alone,), _, _ = _run(pxl, [(f, n)])
        assert np.array_equal(alone, parts[k]), f["name"]


def test_many_small_requests_complete_with_stats(pxl):
    reqs = [(random_request(i, sizes=(16, 24, 32)), 2) for i in range(6)]
    with _server(pxl) as srv:
        ids = [srv.submit(f, FPS, seed=0, steps=n) for f, n in reqs]
        srv.drain(timeout=300)
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0
        assert st["p50_ms"] > 0 and st["p99_ms"] >= st["p50_ms"]
        assert st["particle_steps"] > 0 and st["throughput"] > 0
        for i in ids:
            info = srv.poll(i)
            assert info["status"] == "done" and info["rel_residual"] <= 1e-6


def test_budget_serialises_admission(pxl):
    f = dam_break(24)
    with _server(pxl) as probe:
        rid = probe.submit(f, FPS, steps=0)
        probe.wait(rid, timeout=60)
        one = probe.poll(rid)["device_bytes"]
    with _server(pxl, budget_bytes=one) as srv:
        ids = [srv.submit(f, FPS, seed=s, steps=2) for s in range(3)]
        srv.drain(timeout=300)
        st = srv.stats()
        assert st["completed"] == 3 and st["peak_bytes"] <= one
        assert abs(st["mean_batch"] - 1.0) < 1e-9  # one at a time
        for i in ids:
            assert srv.poll(i)["status"] == "done"
    with _server(pxl, budget_bytes=one - 1) as srv:
        with pytest.raises(pxl.FluidError) as e:
            srv.submit(f, FPS, steps=1)
        assert "budget" in str(e.value)


def test_late_arrival_joins_and_finishes_first(pxl):
    long_f, short_f = dam_break(32), two_blocks(16)
    (p_short_alone,), _, _ = _run(pxl, [(short_f, 2)])
    with _server(pxl) as srv:
        a = srv.submit(long_f, FPS, steps=40)
        t0 = time.time()
        while srv.poll(a)["steps_done"] < 2 and time.time() - t0 < 60:
            time.sleep(0.005)
        assert srv.poll(a)["status"] == "running"
        b = srv.submit(short_f, FPS, steps=2, keep_particles=True)
        ib = srv.wait(b, timeout=120)
        assert ib["status"] == "done"
        ia = srv.poll(a)
        assert ia["status"] == "running" and ia["steps_done"] < 40
        assert np.array_equal(srv.read_particles(b), p_short_alone)
        srv.wait(a, timeout=300)
        assert srv.stats()["mean_batch"] > 1.0


def test_cancel_and_invalid_request(pxl):
    with _server(pxl) as srv:
        a = srv.submit(dam_break(32), FPS, steps=100000)
        t0 = time.time()
        while srv.poll(a)["steps_done"] < 1 and time.time() - t0 < 60:
            time.sleep(0.005)
        srv.cancel(a)
        info = srv.wait(a, timeout=60)
        assert info["status"] == "cancelled"
        with pytest.raises(pxl.FluidError) as e:
            srv.submit({"name": "bad", "resolution": 8}, FPS, steps=1)
        assert "missing field `bounds`" in str(e.value)
        with pytest.raises(pxl.FluidError) as e:
            srv.submit(dict(dam_break(16), cfl=0.5), FPS, steps=1)
        assert "not supported in batched execution" in str(e.value)
        assert srv.stats()["cancelled"] == 1


def test_fixed_batch_matches_server_and_prices_launches(pxl):
    fl = [dam_break(24), two_blocks(16)]
    with pxl.Batch(fl, FPS, solver_workgroups=WG) as b:
        b.step(2)
        pb = [b.read_particles(i) for i in range(2)]
        for i in range(2):
            assert 0 <= b.last_solve(i)["rel_residual"] <= 1e-6
        prof = b.profile_steps(1)
    parts, _, _ = _run(pxl, [(fl[0], 2), (fl[1], 2)])
    assert np.array_equal(pb[0], parts[0]) and np.array_equal(pb[1], parts[1])
    by = {k["kernel"]: k for k in prof}
    for name in ("p2g", "g2p", "faces", "mg_sweep", "mgcg_update", "mg_tail"):
        assert by[name]["launches"] > 0 and by[name]["total_ms"] > 0 and by[name]["alg_bytes"] > 0, name


def test_batch_grids_change_speed_not_results(pxl, tmp_path):
    """The batched launches' persistent-grid maxima only change how work is spread: a batch
    stepped with every family at the minimum (8 workgroups) is bitwise the batch stepped with
    the defaults.  The autotuner then picks grids by timing and writes its JSON cache."""
    fluids = [random_request(i, sizes=(32, 48)) for i in range(3)]
    saved = pxl.batch_grids()
    with pxl.Batch(fluids, FPS) as b:
        b.step(2)
        ref = [b.read_particles(i) for i in range(len(fluids))]
    try:
        pxl.set_batch_grids({k: 8 for k in pxl.GRID_FAMILIES})
        with pxl.Batch(fluids, FPS) as b:
            b.step(2)
            got = [b.read_particles(i) for i in range(len(fluids))]
        for i, (a, c) in enumerate(zip(ref, got)):
            assert np.array_equal(a.view(np.uint32), c.view(np.uint32)), f"member {i} differs with 8-workgroup grids"
        pxl.set_batch_grids(saved)
        cache = str(tmp_path / "tuned.json")
        r = pxl.autotune_batch(fluids, FPS, steps=1, warmup=1, families=("tile", "mg_list"), cache=cache)
        print(f"autotune {r['grids']} {r['ms_per_step']:.3f} ms/step (baseline {r['baseline_ms_per_step']:.3f})",
              flush=True)
        assert set(r["grids"]) == set(pxl.GRID_FAMILIES)
        assert all(v % 8 == 0 and 8 <= v <= 16384 for v in r["grids"].values())
        assert len(r["trials"]) == 5 and r["ms_per_step"] <= r["baseline_ms_per_step"]
        assert pxl.batch_grids() == r["grids"]
        pxl.set_batch_grids(saved)
        assert pxl.load_batch_grids(cache) == r["grids"]
    finally:
        pxl.set_batch_grids(saved)


def test_batch_workgroup_sizes_change_speed_not_results(pxl, tmp_path):
    """The per-tile P2G / G2P kernels in 512-lane workgroups and the band passes in 512-lane
    workgroups step a batch bitwise like the default sizes (1024 and 256); the autotuner times
    the other sizes and records its choice in the JSON cache."""
    fluids = [random_request(i, sizes=(32, 48, 64)) for i in range(4)]
    saved = pxl.batch_shapes()
    with pxl.Batch(fluids, FPS) as b:
        b.step(3)
        ref = [b.read_particles(i) for i in range(len(fluids))]
        ref_p = [b.read_field(i, "pressure64") for i in range(len(fluids))]
    try:
        for shapes in ({"tile_lanes": 512}, {"band_lanes": 512}, {"tile_lanes": 512, "band_lanes": 512}):
            pxl.set_batch_shapes(dict(saved, **shapes))
            with pxl.Batch(fluids, FPS) as b:
                b.step(3)
                for i in range(len(fluids)):
                    assert np.array_equal(ref[i].view(np.uint32), b.read_particles(i).view(np.uint32)), (shapes, i)
                    assert np.array_equal(ref_p[i].view(np.uint64), b.read_field(i, "pressure64").view(np.uint64))
        pxl.set_batch_shapes(saved)
        cache = str(tmp_path / "shapes.json")
        r = pxl.autotune_batch(fluids, FPS, steps=1, warmup=1, families=("tile_lanes", "band_lanes"), cache=cache)
        print(f"shape autotune {r['shapes']} {r['ms_per_step']:.3f} ms/step", flush=True)
        assert len(r["trials"]) == 3 and set(r["shapes"]) == {"tile_lanes", "band_lanes"}
        assert pxl.batch_shapes() == r["shapes"]
        import json
        assert json.load(open(cache))["shapes"] == r["shapes"]
    finally:
        pxl.set_batch_shapes(saved)
