"""GPU agent neighbour grid (shaders/sim.slang:27-104) against the oracle's stable counting
sort on the same agents: counts, offsets and the sorted ids bit for bit (ids ascending within
a cell, the order the radix sort produces), for crowded cells, negative cells, tile
boundaries, every radix pass count and the empty set."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pxl():
    import pxlengine_amd
    return pxlengine_amd


def _agents(n, seed, crowd=0.25, spread=6.0):
    rng = np.random.default_rng(seed)
    a = np.zeros((n, 16), np.float32)
    a[:, 0:3] = rng.normal(0.0, spread, (n, 3)).astype(np.float32)
    a[: int(n * crowd), 0:3] = np.float32([0.3, 0.3, 0.3])
    a[:, 3:] = rng.uniform(-1, 1, (n, 13)).astype(np.float32)
    return a


@pytest.mark.parametrize("n,cell,table", [
    (1, 1.0, None), (5000, 0.7, None), (2049, 1.6, None), (100_000, 1.6, None),
    (30_000, 0.5, 1 << 23), (4096, 2.0, 1 << 8), (0, 1.0, None),
], ids=["one", "crowded", "tile+1", "100k", "three-passes", "one-pass", "empty"])
def test_grid_bitwise(pxl, orc, n, cell, table):
    a = _agents(n, seed=n + 7)
    t = pxl.AgentGrid.table_size(n) if table is None else table
    g = pxl.AgentGrid(max(n, 1))
    off, cnt, srt = g.build(a, cell, t)
    o_off, o_cnt, o_srt = orc.grid_build(a, cell, t)
    np.testing.assert_array_equal(cnt, o_cnt)
    np.testing.assert_array_equal(off, o_off)
    np.testing.assert_array_equal(srt, o_srt)
    g.close()


def test_grid_device_buffers_and_errors(pxl, orc):
    n = 50_000
    a = _agents(n, seed=3)
    t = pxl.AgentGrid.table_size(n)
    g = pxl.AgentGrid(n)
    bufs = [pxl.DeviceBuffer(b) for b in (a.nbytes, 4 * t, 4 * t, 4 * n)]
    try:
        bufs[0].upload(a)
        ms = g.build_device(bufs[0].ptr, n, 1.6, t, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr)
        assert ms > 0
        off = bufs[1].download(np.zeros(t, np.uint32))
        cnt = bufs[2].download(np.zeros(t, np.uint32))
        srt = bufs[3].download(np.zeros(n, np.uint32))
    finally:
        for b in bufs:
            b.free()
    o_off, o_cnt, o_srt = orc.grid_build(a, 1.6, t)
    np.testing.assert_array_equal(off, o_off)
    np.testing.assert_array_equal(cnt, o_cnt)
    np.testing.assert_array_equal(srt, o_srt)
    with pytest.raises(pxl.FluidError, match="exceeds the grid's max_agents"):
        g.build(_agents(n + 1, seed=4), 1.6)
    with pytest.raises(pxl.FluidError, match="power of two"):
        g.build(a, 1.6, 3000)
    g.close()
