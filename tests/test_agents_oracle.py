"""Pins the oracle's agent neighbour grid (oracle/flip_oracle.c, restating
shaders/sim.slang:27-104 and crates/pxl-render/src/sim.rs:129, 178-186): the hash in exact
Python integers (wrapping u32 products, negative cells as two's complement), the table size
closed form, and the grid against numpy's stable sort."""
import numpy as np
import pytest


def _hash(c, t):
    x, y, z = (int(v) & 0xFFFFFFFF for v in c)
    return (((x * 73856093) & 0xFFFFFFFF) ^ ((y * 19349663) & 0xFFFFFFFF) ^ ((z * 83492791) & 0xFFFFFFFF)) & (t - 1)


def test_table_size(orc):
    assert orc.grid_table_size(0) == 1024 and orc.grid_table_size(1) == 1024
    assert orc.grid_table_size(512) == 1024 and orc.grid_table_size(513) == 2048
    assert orc.grid_table_size(1_000_000) == 1 << 21  # the flock example: 2M -> 2^21
    assert orc.grid_table_size(10_000_000) == 1 << 23  # clamp


def test_cell_hash_exact_integers(orc):
    rng = np.random.default_rng(1)
    for c in rng.integers(-50_000, 50_000, size=(500, 3)):
        for t in (1024, 1 << 21):
            assert orc.cell_hash(*(int(v) for v in c), t) == _hash(c, t)


@pytest.mark.parametrize("n,cell", [(1, 1.0), (5000, 0.7), (20000, 1.6)])
def test_grid_is_the_stable_counting_sort(orc, n, cell):
    rng = np.random.default_rng(n)
    agents = np.zeros((n, 16), np.float32)
    agents[:, 0:3] = rng.normal(0.0, 6.0, (n, 3)).astype(np.float32)  # negative cells too
    agents[: n // 4, 0:3] = np.float32([0.3, 0.3, 0.3])                 # one crowded cell
    agents[:, 3:] = rng.uniform(-1, 1, (n, 13)).astype(np.float32)     # other fields ignored
    t = orc.grid_table_size(n)
    off, cnt, srt = orc.grid_build(agents, cell)
    cells = np.floor(agents[:, 0:3] / np.float32(cell)).astype(np.int64)
    keys = np.array([_hash(c, t) for c in cells], np.int64)
    np.testing.assert_array_equal(cnt, np.bincount(keys, minlength=t))
    np.testing.assert_array_equal(off, np.concatenate([[0], np.cumsum(cnt)[:-1]]))
    np.testing.assert_array_equal(srt, np.argsort(keys, kind="stable"))
