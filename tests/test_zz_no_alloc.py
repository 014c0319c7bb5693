"""pxlgrid_build_static -- the agent neighbour grid without allocation, its scratch a static
device array -- called through the C ABI gives bitwise the outputs of pxlgrid_build (a handle
with its own allocated scratch), on two calls in a row, the second with fewer agents, a smaller
table and output buffers pre-filled with garbage."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _flock(n, seed):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-60.0, 60.0, (8, 3))
    spreads = rng.uniform(2.0, 12.0, 8)
    which = rng.integers(0, 8, n)
    a = np.zeros((n, 16), np.float32)
    a[:, 0:3] = (centres[which] + rng.normal(0.0, 1.0, (n, 3)) * spreads[which, None]).astype(np.float32)
    return a


def test_static_build_equals_handle_build():
    import pxlengine_amd as pxl
    lib = pxl.load_library()
    cases = [(300_000, 1.6, 0), (70_000, 0.9, 1)]
    g = pxl.AgentGrid(cases[0][0])
    for n, cell, seed in cases:
        print(f"static grid build: {n} agents, cell {cell}", flush=True)
        a = _flock(n, seed)
        t = pxl.AgentGrid.table_size(n)
        agents = pxl.DeviceBuffer(a.nbytes)
        agents.upload(a)
        ref = [pxl.DeviceBuffer(b) for b in (4 * t, 4 * t, 4 * n)]
        out = [pxl.DeviceBuffer(b) for b in (4 * t, 4 * t, 4 * n)]
        for b, m in zip(out, (t, t, n)):
            b.upload(np.full(m, 0xABABABAB, np.uint32))
        g.build_device(agents.ptr, n, cell, t, *(b.ptr for b in ref))
        rc = lib.pxlgrid_build_static(0, None, C.c_void_p(agents.ptr), n, float(np.float32(cell)), t,
                                      *(C.c_void_p(b.ptr) for b in out))
        assert rc == 0, lib.pxlgrid_last_error(None).decode()
        for name, r, o, m in zip(("offsets", "counts", "sorted"), ref, out, (t, t, n)):
            exp = r.download(np.zeros(m, np.uint32))
            got = o.download(np.zeros(m, np.uint32))
            np.testing.assert_array_equal(got, exp, err_msg=f"{n} agents: {name}")
        assert int(exp.size) == n and int(ref[1].download(np.zeros(t, np.uint32)).sum()) == n
        for b in [agents] + ref + out:
            b.free()
    g.close()
    too_many = (1 << 20) + 1  # PXLGRID_STATIC_MAX_AGENTS + 1
    rc = lib.pxlgrid_build_static(0, None, C.c_void_p(256), too_many, 1.0, 1024, C.c_void_p(256), C.c_void_p(256),
                                  C.c_void_p(256))
    assert rc == pxl.flip.ERR_INVALID and "exceeds the static scratch" in lib.pxlgrid_last_error(None).decode()
