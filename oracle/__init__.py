"""CPU oracle of pxlengine's FLIP step -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/liborc_flip.so`` (built from ``flip_oracle.c`` by ``oracle/Makefile``) and
wraps it with numpy views.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
cpu_baseline leg may import this module, and only as the checker; the product package
``pxlengine_amd`` never imports it.

The restated reference semantics live in ``flip_oracle.c``; every function there cites
``shaders/fluid.slang`` / ``crates/pxl-render/src/fluid.rs`` / ``crates/pxl-scene/src/lib.rs``
line ranges.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc_flip.so")

AIR, FLUID, SOLID = 0, 1, 2


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32), ("n_particles", C.c_uint32),
        ("h", C.c_float), ("dt", C.c_float), ("flip", C.c_float), ("inv_h", C.c_float),
        ("origin", C.c_float * 3), ("gravity", C.c_float * 3),
    ]


class _State(C.Structure):
    _fields_ = [
        ("p", Params), ("n_faces", C.c_uint64), ("n_cells", C.c_uint64),
        ("px", C.POINTER(C.c_float)), ("py", C.POINTER(C.c_float)), ("pz", C.POINTER(C.c_float)),
        ("vx", C.POINTER(C.c_float)), ("vy", C.POINTER(C.c_float)), ("vz", C.POINTER(C.c_float)),
        ("vel", C.POINTER(C.c_float)), ("vel_old", C.POINTER(C.c_float)),
        ("acc", C.POINTER(C.c_int32)), ("wacc", C.POINTER(C.c_int32)),
        ("count", C.POINTER(C.c_uint32)), ("type", C.POINTER(C.c_uint32)),
        ("pressure", C.POINTER(C.c_float)), ("div", C.POINTER(C.c_float)),
        ("obst_id", C.POINTER(C.c_uint32)), ("obst", C.POINTER(C.c_float)), ("n_obst", C.c_uint32),
    ]


class SurfaceParams(C.Structure):
    _fields_ = [
        ("sx", C.c_uint32), ("sy", C.c_uint32), ("sz", C.c_uint32),
        ("max_verts", C.c_uint32), ("max_tris", C.c_uint32),
        ("sh", C.c_float), ("radius", C.c_float), ("iso", C.c_float), ("origin", C.c_float * 3),
    ]

    @property
    def nodes(self):
        return (self.sx + 1) * (self.sy + 1) * (self.sz + 1)

    @property
    def cells(self):
        return self.sx * self.sy * self.sz


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    cmd = ["make", "-s", "-C", _HERE] + (["-B"] if force else [])
    subprocess.run(cmd, check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        f32p, u32p, dp = C.POINTER(C.c_float), C.POINTER(C.c_uint32), C.POINTER(C.c_double)
        SP = C.POINTER(_State)
        L.orc_fluid_grid.argtypes = [f32p, f32p, C.c_uint32, u32p, f32p]
        L.orc_lcg.argtypes = [C.POINTER(C.c_uint64)]
        L.orc_lcg.restype = C.c_float
        L.orc_seed.argtypes = [u32p, C.c_float, f32p, f32p, C.c_uint32, C.c_uint32, C.c_uint32, u32p, f32p, f32p]
        L.orc_seed.restype = C.c_uint64
        L.orc_obstacle_cells.argtypes = [u32p, C.c_float, f32p, f32p, C.c_uint32, u32p]
        L.orc_obstacle_cells.restype = C.c_uint64
        L.orc_set_obstacles.argtypes = [SP, f32p, C.c_uint32]
        L.orc_set_obstacles.restype = C.c_int
        L.orc_steps_for_frame.argtypes = [C.c_double, C.c_double, C.c_double, C.c_uint32]
        L.orc_steps_for_frame.restype = C.c_uint64
        L.orc_dt.argtypes = [C.c_double, C.c_uint32]
        L.orc_dt.restype = C.c_float
        L.orc_create.argtypes = [C.POINTER(Params)]
        L.orc_create.restype = SP
        L.orc_destroy.argtypes = [SP]
        for name in ("orc_clear_grid", "orc_p2g", "orc_normalize_grid", "orc_forces",
                     "orc_divergence", "orc_project", "orc_g2p_advect"):
            getattr(L, name).argtypes = [SP]
        L.orc_pressure_rb.argtypes = [SP, C.c_uint32]
        L.orc_step_rbgs.argtypes = [SP, C.c_uint32]
        L.orc_step_cg.argtypes = [SP, C.c_uint32]
        L.orc_pressure_cg32.argtypes = [SP, C.c_uint32]
        L.orc_bin.argtypes = [SP, u32p, u32p, u32p, u32p]
        L.orc_poisson_apply64.argtypes = [SP, dp, dp]
        L.orc_poisson_rhs64.argtypes = [SP, dp]
        L.orc_poisson_relres64.argtypes = [SP, f32p, dp]
        L.orc_poisson_relres64.restype = C.c_double
        L.orc_poisson_solve64.argtypes = [SP, dp, C.c_double, C.c_int, dp]
        L.orc_poisson_solve64.restype = C.c_int
        L.orc_grid_table_size.argtypes = [C.c_uint32]
        L.orc_grid_table_size.restype = C.c_uint32
        L.orc_cell_hash.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_uint32]
        L.orc_cell_hash.restype = C.c_uint32
        L.orc_grid_build.argtypes = [f32p, C.c_uint32, C.c_float, C.c_uint32, u32p, u32p, u32p]
        L.orc_cfl_substeps.argtypes = [SP, C.c_float, C.c_float]
        L.orc_cfl_substeps.restype = C.c_uint32
        SPP = C.POINTER(SurfaceParams)
        L.orc_surface_params.argtypes = [u32p, C.c_float, f32p, C.c_uint32, C.c_float, C.c_uint32, C.c_uint32, SPP]
        L.orc_surface_splat.argtypes = [SPP, f32p, f32p, f32p, C.c_uint32, C.POINTER(C.c_int32)]
        L.orc_surface_phi.argtypes = [SPP, C.POINTER(C.c_int32), f32p]
        L.orc_surface_mesh.argtypes = [SPP, f32p, u32p, f32p, u32p, u32p]
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _up(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def fluid_grid(lo, hi, resolution):
    """Fluid::grid (crates/pxl-scene/src/lib.rs:123-128) -> ((nx, ny, nz), h)."""
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    n = np.zeros(3, np.uint32)
    h = C.c_float()
    lib().orc_fluid_grid(_fp(lo), _fp(hi), resolution, _up(n), C.byref(h))
    return tuple(int(v) for v in n), np.float32(h.value)


def lcg_stream(state: int, count: int):
    s = C.c_uint64(state)
    out = np.empty(count, np.float32)
    for i in range(count):
        out[i] = lib().orc_lcg(C.byref(s))
    return out, s.value


def _obstacle_rows(obstacles):
    return np.asarray([list(c) + [r] for c, r in obstacles], np.float32).reshape(-1)


def obstacle_cells(n, h, origin, obstacles):
    """Obstacle cells (an extension): per cell 1 + the index of the first obstacle (list order)
    holding its centre, 0 elsewhere.  obstacles: list of (centre3, radius)."""
    nn = np.asarray(n, np.uint32)
    org = np.asarray(origin, np.float32)
    ob = _obstacle_rows(obstacles)
    out = np.zeros(int(nn[0]) * int(nn[1]) * int(nn[2]), np.uint32)
    lib().orc_obstacle_cells(_up(nn), np.float32(h), _fp(org), _fp(ob), len(obstacles), _up(out))
    return out


def seed_particles(n, h, origin, blocks, ppc, seed, count_only=False, obstacles=None):
    """Seeding loop of FluidState::new (crates/pxl-render/src/fluid.rs:123-148).

    blocks: list of (min3, max3, velocity3) boxes, or ("sphere", centre3, radius, velocity3)
    sphere fills (an extension).  obstacles (an extension): list of (centre3, radius) solid
    spheres whose cells receive no particles.  Returns (x[P,3], v[P,3]) float32, or the count.
    """
    nn = np.asarray(n, np.uint32)
    org = np.asarray(origin, np.float32)
    rows = []
    for blk in blocks:
        if len(blk) == 4 and blk[0] == "sphere":
            _, c, r, v = blk
            rows.append([1.0] + list(c) + [r, 0.0, 0.0] + list(v))
        else:
            m, M, v = blk
            rows.append([0.0] + list(m) + list(M) + list(v))
    b = np.asarray(rows, np.float32).reshape(-1)
    L = lib()
    ids = obstacle_cells(n, h, origin, obstacles) if obstacles else None
    idp = _up(ids) if ids is not None else None
    cnt = L.orc_seed(_up(nn), np.float32(h), _fp(org), _fp(b), len(blocks), ppc, seed, idp, None, None)
    if count_only:
        return int(cnt)
    x = np.zeros((max(cnt, 1), 3), np.float32)
    v = np.zeros((max(cnt, 1), 3), np.float32)
    L.orc_seed(_up(nn), np.float32(h), _fp(org), _fp(b), len(blocks), ppc, seed, idp, _fp(x), _fp(v))
    return x[:cnt], v[:cnt]


def surface_params(n, h, origin, refine=2, particle_radius=2.0, ppc=8, max_triangles=2_000_000):
    """Surface-net parameters of a fluid (crates/pxl-render/src/fluid.rs:151-157)."""
    nn = np.asarray(n, np.uint32)
    org = np.asarray(origin, np.float32)
    sp = SurfaceParams()
    lib().orc_surface_params(_up(nn), np.float32(h), _fp(org), refine, np.float32(particle_radius), ppc,
                             max_triangles, C.byref(sp))
    return sp


def surface_splat(sp, x):
    """s_clear + s_splat: fixed-point node density of the particles x[P, 3]."""
    x = np.ascontiguousarray(x, np.float32)
    cols = [np.ascontiguousarray(x[:, a]) for a in range(3)]
    d = np.zeros(sp.nodes, np.int32)
    lib().orc_surface_splat(C.byref(sp), _fp(cols[0]), _fp(cols[1]), _fp(cols[2]), len(x),
                            d.ctypes.data_as(C.POINTER(C.c_int32)))
    return d


def surface_phi(sp, density):
    d = np.ascontiguousarray(density, np.int32)
    phi = np.zeros(sp.nodes, np.float32)
    lib().orc_surface_phi(C.byref(sp), d.ctypes.data_as(C.POINTER(C.c_int32)), _fp(phi))
    return phi


def surface_mesh(sp, phi):
    """s_verts + s_quads in index order -> (cell_vert, verts[V, 8], indices[T, 3], (vertices,
    triangles) counters before clipping)."""
    phi = np.ascontiguousarray(phi, np.float32)
    cv = np.zeros(sp.cells, np.uint32)
    verts = np.zeros((max(sp.max_verts, 1), 8), np.float32)
    idx = np.zeros((max(sp.max_tris, 1), 3), np.uint32)
    ctr = np.zeros(2, np.uint32)
    lib().orc_surface_mesh(C.byref(sp), _fp(phi), _up(cv), _fp(verts), _up(idx), _up(ctr))
    nv = min(int(ctr[0]), sp.max_verts)
    nt = min(int(ctr[1]), sp.max_tris) // 2 * 2
    return cv, verts[:nv], idx[:nt], (int(ctr[0]), int(ctr[1]))


def grid_table_size(count):
    return int(lib().orc_grid_table_size(count))


def cell_hash(cx, cy, cz, table_size):
    return int(lib().orc_cell_hash(cx, cy, cz, table_size))


def grid_build(agents, cell_size, table_size=None):
    """Agent neighbour grid of agents[N, 16] (64-byte records) -> (offsets, counts, sorted)."""
    a = np.ascontiguousarray(agents, np.float32)
    n = len(a)
    t = grid_table_size(n) if table_size is None else table_size
    off = np.zeros(t, np.uint32)
    cnt = np.zeros(t, np.uint32)
    srt = np.zeros(max(n, 1), np.uint32)
    lib().orc_grid_build(_fp(a), n, np.float32(cell_size), t, _up(off), _up(cnt), _up(srt))
    return off, cnt, srt[:n]


def steps_for_frame(warmup, frame, fps, substeps):
    return int(lib().orc_steps_for_frame(float(np.float32(warmup)), frame, fps, substeps))


def dt_for(fps, substeps):
    return np.float32(lib().orc_dt(fps, substeps))


class Sim:
    """One oracle simulation; arrays are numpy views of the C state (SoA particles)."""

    def __init__(self, n, h, origin, dt, flip, gravity, x, v, obstacles=None):
        p = Params()
        p.nx, p.ny, p.nz = (int(a) for a in n)
        p.n_particles = len(x)
        p.h = float(np.float32(h))
        p.dt = float(np.float32(dt))
        p.flip = float(np.float32(flip))
        p.inv_h = float(np.float32(1.0) / np.float32(h))
        for a in range(3):
            p.origin[a] = float(np.float32(origin[a]))
            p.gravity[a] = float(np.float32(gravity[a]))
        self._L = lib()
        self._s = self._L.orc_create(C.byref(p))
        if not self._s:
            raise MemoryError("oracle allocation failed")
        st = self._s.contents
        self.n = tuple(int(a) for a in n)
        self.P = len(x)
        self.F = int(st.n_faces)
        self.C = int(st.n_cells)

        def view(ptr, count, dtype):
            return np.ctypeslib.as_array(ptr, shape=(max(count, 1),))[:count].view(dtype)

        self.px, self.py, self.pz = (view(getattr(st, k), self.P, np.float32) for k in ("px", "py", "pz"))
        self.vx, self.vy, self.vz = (view(getattr(st, k), self.P, np.float32) for k in ("vx", "vy", "vz"))
        self.vel = view(st.vel, self.F, np.float32)
        self.vel_old = view(st.vel_old, self.F, np.float32)
        self.acc = view(st.acc, self.F, np.int32)
        self.wacc = view(st.wacc, self.F, np.int32)
        self.count = view(st.count, self.C, np.uint32)
        self.type = view(st.type, self.C, np.uint32)
        self.pressure = view(st.pressure, self.C, np.float32)
        self.div = view(st.div, self.C, np.float32)
        x = np.asarray(x, np.float32)
        v = np.asarray(v, np.float32)
        self.px[:], self.py[:], self.pz[:] = x[:, 0], x[:, 1], x[:, 2]
        self.vx[:], self.vy[:], self.vz[:] = v[:, 0], v[:, 1], v[:, 2]
        self.params = p
        self.obst_id = None
        if obstacles:
            ob = _obstacle_rows(obstacles)
            if self._L.orc_set_obstacles(self._s, _fp(ob), len(obstacles)) != 0:
                raise MemoryError("oracle allocation failed")
            self.obst_id = view(self._s.contents.obst_id, self.C, np.uint32)

    def __del__(self):
        s = getattr(self, "_s", None)
        if s:
            self._L.orc_destroy(s)
            self._s = None

    # -- particle access
    @property
    def x(self):
        return np.stack([self.px, self.py, self.pz], axis=1)

    @property
    def v(self):
        return np.stack([self.vx, self.vy, self.vz], axis=1)

    def set_particles(self, x, v):
        x = np.asarray(x, np.float32)
        v = np.asarray(v, np.float32)
        self.px[:], self.py[:], self.pz[:] = x[:, 0], x[:, 1], x[:, 2]
        self.vx[:], self.vy[:], self.vz[:] = v[:, 0], v[:, 1], v[:, 2]

    # -- CFL cap (extension)
    def cfl_substeps(self, cfl):
        return int(self._L.orc_cfl_substeps(self._s, np.float32(self.params.dt), np.float32(cfl)))

    def set_dt(self, dt):
        self._s.contents.p.dt = float(np.float32(dt))
        self.params.dt = float(np.float32(dt))

    def step_rbgs_capped(self, iters, cfl):
        """One step of dt under the CFL cap: k = cfl_substeps, then k reference steps of dt / k."""
        dt = np.float32(self.params.dt)
        k = self.cfl_substeps(cfl)
        if k > 1:
            self.set_dt(dt / np.float32(k))
        try:
            for _ in range(k):
                self.step_rbgs(iters)
        finally:
            self.set_dt(dt)
        return k

    # -- kernels
    def clear_grid(self): self._L.orc_clear_grid(self._s)
    def p2g(self): self._L.orc_p2g(self._s)
    def normalize_grid(self): self._L.orc_normalize_grid(self._s)
    def forces(self): self._L.orc_forces(self._s)
    def divergence(self): self._L.orc_divergence(self._s)
    def pressure_rb(self, parity): self._L.orc_pressure_rb(self._s, parity)
    def project(self): self._L.orc_project(self._s)
    def g2p_advect(self): self._L.orc_g2p_advect(self._s)
    def step_rbgs(self, iters): self._L.orc_step_rbgs(self._s, iters)
    def step_cg(self, iters): self._L.orc_step_cg(self._s, iters)
    def pressure_cg32(self, iters): self._L.orc_pressure_cg32(self._s, iters)

    def grid_phase(self):
        """clear -> p2g -> normalize -> forces -> divergence (fluid.rs:290-294)."""
        self.clear_grid()
        self.p2g()
        self.normalize_grid()
        self.forces()
        self.divergence()

    def bin(self):
        cell_of = np.zeros(max(self.P, 1), np.uint32)
        counts = np.zeros(self.C, np.uint32)
        offsets = np.zeros(self.C, np.uint32)
        ids = np.zeros(max(self.P, 1), np.uint32)
        self._L.orc_bin(self._s, _up(cell_of), _up(counts), _up(offsets), _up(ids))
        return cell_of[:self.P], counts, offsets, ids[:self.P]

    # -- FP64 Poisson oracle
    def poisson_apply64(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.C, np.float64)
        self._L.orc_poisson_apply64(self._s, _dp(x), _dp(y))
        return y

    def poisson_rhs64(self):
        b = np.zeros(self.C, np.float64)
        self._L.orc_poisson_rhs64(self._s, _dp(b))
        return b

    def poisson_relres64(self, x):
        x = np.asarray(x)
        if x.dtype == np.float64:
            x = np.ascontiguousarray(x)
            return float(self._L.orc_poisson_relres64(self._s, None, _dp(x)))
        x = np.ascontiguousarray(x, np.float32)
        return float(self._L.orc_poisson_relres64(self._s, _fp(x), None))

    def poisson_solve64(self, tol=1e-12, max_iter=100000):
        x = np.zeros(self.C, np.float64)
        rr = C.c_double()
        it = self._L.orc_poisson_solve64(self._s, _dp(x), tol, max_iter, C.byref(rr))
        return x, int(it), float(rr.value)
