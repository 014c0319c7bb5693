"""Pins the CPU oracle (oracle/flip_oracle.c) to facts that do not come from the oracle.

Sources of truth, per function:
  grid       the reference's own test fluids_validate (crates/pxl-scene/src/lib.rs:886-899)
             and closed forms of Fluid::grid;
  seeding    ROADMAP's published particle count for examples/dam-break.json (139k,
             docs/ROADMAP.md:83), the LCG recurrence in exact integer arithmetic, and
             positions recomputed in numpy float32 from that integer stream;
  binning    the textbook counting sort (numpy histogram / cumsum / stable argsort);
  P2G        partition of unity and reproduction of the particle position by the weighted
             face positions (first moment of trilinear weights);
  grid ops   closed forms (gravity adds g*dt; linear field -> constant divergence);
  pressure   the hydrostatic known answer p_j = |g| h (n - j + 1) and a scipy sparse
             assembly of the same matrix;
  G2P        uniform fields (PIC = field, FLIP delta = 0), free fall with flip = 1,
             wall clamping;
  scheduling steps_for_frame / dt closed forms.
"""
import numpy as np
import pytest

from tests.scenes import EXAMPLE_DAM_BREAK, FPS, dam_break, obstacle_scene, oracle_sim

G = np.float32(9.81)


# ---------------------------------------------------------------- grid / seeding / schedule

def test_fluid_grid_reference_test(orc):
    # fluids_validate: bounds [[0,0,0],[2,1,1]], resolution 32 -> [32,16,16], h = 0.0625
    n, h = orc.fluid_grid([0, 0, 0], [2, 1, 1], 32)
    assert n == (32, 16, 16)
    assert abs(h - 0.0625) < 1e-6


def test_fluid_grid_closed_forms(orc):
    # dam-break example: size (2, 1.2, 1), h = 2/64, n = round(size/h) = (64, 38.4->38, 32)
    n, h = orc.fluid_grid(*EXAMPLE_DAM_BREAK["bounds"], 64)
    assert n == (64, 38, 32) and h == np.float32(0.03125)
    # tiny axis is lifted to 4 cells; rounding is half away from zero (1.5 -> 2)
    n, h = orc.fluid_grid([0, 0, 0], [1, 0.01, 1], 16)
    assert n == (16, 4, 16)
    n, h = orc.fluid_grid([0, 0, 0], [16, 1.5, 1], 16)
    assert n[1] == 2 or n[1] == 4  # 1.5 -> 2, then max(4)
    assert n == (16, 4, 4)
    n, h = orc.fluid_grid([0, 0, 0], [32, 10, 1], 64)  # h = 0.5: 10/0.5 = 20
    assert n == (64, 20, 4) and h == np.float32(0.5)


def _lcg_ints(state, count):
    out = []
    for _ in range(count):
        state = (state * 6364136223846793005 + 1442695040888963407) % (1 << 64)
        out.append(state >> 40)
    return out, state


def test_lcg_closed_form(orc):
    for seed in (0, 1, 12345):
        s0 = 0x9E3779B97F4A7C15 ^ seed
        ints, s_end = _lcg_ints(s0, 50)
        vals, s = orc.lcg_stream(s0, 50)
        assert s == s_end
        np.testing.assert_array_equal(vals, np.array(ints, np.float64).astype(np.float32) / np.float32(1 << 24))


def test_seed_count_published(orc):
    # docs/ROADMAP.md:83: "Dam break: 139k particles, 64-cell grid"
    f = EXAMPLE_DAM_BREAK
    n, h = orc.fluid_grid(*f["bounds"], 64)
    cnt = orc.seed_particles(n, h, f["bounds"][0], [(b["min"], b["max"], [0, 0, 0]) for b in f["fill"]], 8, 0,
                             count_only=True)
    assert cnt == 139200


def _cells_in_block(n, h, origin, bmin, bmax):
    """Interior cells (1..n-2) whose f32 centre lies in the block, enumerated k, j, i."""
    out = []
    o = np.asarray(origin, np.float32)
    for k in range(1, n[2] - 1):
        for j in range(1, n[1] - 1):
            for i in range(1, n[0] - 1):
                c = o + (np.array([i, j, k], np.float32) + np.float32(0.5)) * np.float32(h)
                if np.all(c >= np.asarray(bmin, np.float32)) and np.all(c <= np.asarray(bmax, np.float32)):
                    out.append((i, j, k))
    return out


@pytest.mark.parametrize("res,ppc,seed", [(16, 8, 0), (20, 3, 1), (12, 5, 99)])
def test_seed_positions_closed_form(orc, res, ppc, seed):
    """Positions = origin + (idx + lcg) * h in f32, cells in k, j, i order, LCG stream in
    exact integers; count = ppc * #cells."""
    f = dam_break(res, ppc=ppc)
    lo, hi = f["bounds"]
    n, h = orc.fluid_grid(lo, hi, res)
    blk = f["fill"][0]
    x, v = orc.seed_particles(n, h, lo, [(blk["min"], blk["max"], [0.5, -1, 2])], ppc, seed)
    cells = _cells_in_block(n, h, lo, blk["min"], blk["max"])
    assert len(x) == ppc * len(cells)
    ints, _ = _lcg_ints(0x9E3779B97F4A7C15 ^ seed, 3 * len(x))
    r = (np.array(ints, np.float64).astype(np.float32) / np.float32(1 << 24)).reshape(-1, 3)
    idx = np.repeat(np.array(cells, np.float32), ppc, axis=0)
    expect = np.asarray(lo, np.float32) + (idx + r) * np.float32(h)
    np.testing.assert_array_equal(x, expect)
    np.testing.assert_array_equal(v, np.tile(np.float32([0.5, -1, 2]), (len(x), 1)))
    # every particle inside its seeding cell
    assert np.all(x >= np.asarray(lo, np.float32) + idx * np.float32(h))
    assert np.all(x <= np.asarray(lo, np.float32) + (idx + 1) * np.float32(h))


def _cells_in_sphere(n, h, origin, centre, radius):
    """Interior cells whose f32 centre c has (cx-x0)^2 + (cy-y0)^2 + (cz-z0)^2 <= r^2 in f32."""
    out = []
    o = np.asarray(origin, np.float32)
    c0 = np.asarray(centre, np.float32)
    r = np.float32(radius)
    for k in range(1, n[2] - 1):
        for j in range(1, n[1] - 1):
            for i in range(1, n[0] - 1):
                c = o + (np.array([i, j, k], np.float32) + np.float32(0.5)) * np.float32(h)
                d = c - c0
                if (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2] <= r * r:
                    out.append((i, j, k))
    return out


@pytest.mark.parametrize("res,ppc,seed", [(16, 4, 0), (22, 3, 7)])
def test_seed_sphere_brute_force(orc, res, ppc, seed):
    """Sphere fills (an extension): cells by brute force, positions from the exact LCG
    stream; a box block after the sphere continues the same stream."""
    lo, hi = [0, 0, 0], [1, 1, 1]
    n, h = orc.fluid_grid(lo, hi, res)
    centre, radius = [0.5, 0.45, 0.55], 0.3
    box = ([0.05, 0.05, 0.05], [0.3, 0.2, 0.9])
    x, v = orc.seed_particles(n, h, lo, [("sphere", centre, radius, [1, 0, -1]), (box[0], box[1], [0, 2, 0])],
                              ppc, seed)
    sc = _cells_in_sphere(n, h, lo, centre, radius)
    bc = _cells_in_block(n, h, lo, *box)
    cells = sc + bc
    assert len(x) == ppc * len(cells)
    # the number of cells approximates the sphere's volume
    assert abs(len(sc) * h ** 3 / (4 / 3 * np.pi * radius ** 3) - 1) < 0.1
    ints, _ = _lcg_ints(0x9E3779B97F4A7C15 ^ seed, 3 * len(x))
    r = (np.array(ints, np.float64).astype(np.float32) / np.float32(1 << 24)).reshape(-1, 3)
    idx = np.repeat(np.array(cells, np.float32), ppc, axis=0)
    np.testing.assert_array_equal(x, np.asarray(lo, np.float32) + (idx + r) * np.float32(h))
    ns = ppc * len(sc)
    np.testing.assert_array_equal(v[:ns], np.tile(np.float32([1, 0, -1]), (ns, 1)))
    np.testing.assert_array_equal(v[ns:], np.tile(np.float32([0, 2, 0]), (len(x) - ns, 1)))


def test_obstacle_cells_brute_force(orc):
    """Obstacle cells (an extension): interior cells whose f32 centre a numpy brute force finds
    in a sphere, labelled with the first such sphere in list order (the scene's last sphere
    overlaps its first)."""
    f = obstacle_scene(20)
    lo, hi = f["bounds"]
    n, h = orc.fluid_grid(lo, hi, 20)
    obst = [(o["center"], o["radius"]) for o in f["obstacles"]]
    ids = orc.obstacle_cells(n, h, lo, obst)
    expect = np.zeros(n[0] * n[1] * n[2], np.uint32)
    overlap = 0
    for k in range(1, n[2] - 1):
        for j in range(1, n[1] - 1):
            for i in range(1, n[0] - 1):
                c = np.asarray(lo, np.float32) + (np.array([i, j, k], np.float32) + np.float32(0.5)) * np.float32(h)
                hits = [q for q, (o, r) in enumerate(obst)
                        if (lambda d: (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])(c - np.asarray(o, np.float32))
                        <= np.float32(r) * np.float32(r)]
                if hits:
                    expect[i + n[0] * (j + n[1] * k)] = hits[0] + 1
                    overlap += len(hits) > 1
    np.testing.assert_array_equal(ids, expect)
    assert overlap > 0 and set(np.unique(ids)) == {0, 1, 2, 3, 4}


def test_seed_skips_obstacle_cells(orc):
    """Fill cells that are obstacle cells receive no particles and draw nothing from the LCG:
    the positions are the closed form over the remaining cells."""
    f = obstacle_scene(16, ppc=3)
    lo, hi = f["bounds"]
    n, h = orc.fluid_grid(lo, hi, 16)
    blk = f["fill"][0]
    obst = [(o["center"], o["radius"]) for o in f["obstacles"]]
    ids = orc.obstacle_cells(n, h, lo, obst)
    x, v = orc.seed_particles(n, h, lo, [(blk["min"], blk["max"], [0, 0, 0])], 3, 2, obstacles=obst)
    cells = [c for c in _cells_in_block(n, h, lo, blk["min"], blk["max"])
             if ids[c[0] + n[0] * (c[1] + n[1] * c[2])] == 0]
    assert 0 < len(cells) < len(_cells_in_block(n, h, lo, blk["min"], blk["max"]))
    assert len(x) == 3 * len(cells)
    ints, _ = _lcg_ints(0x9E3779B97F4A7C15 ^ 2, 3 * len(x))
    r = (np.array(ints, np.float64).astype(np.float32) / np.float32(1 << 24)).reshape(-1, 3)
    idx = np.repeat(np.array(cells, np.float32), 3, axis=0)
    np.testing.assert_array_equal(x, np.asarray(lo, np.float32) + (idx + r) * np.float32(h))


def test_seed_counts_configs(orc):
    n, h = orc.fluid_grid([0, 0, 0], [1, 1, 1], 64)
    assert orc.seed_particles(n, h, [0, 0, 0], [([0, 0, 0], [0.4, 0.6, 1], [0, 0, 0])], 8, 0, count_only=True) \
        == 8 * len(_cells_in_block(n, h, [0, 0, 0], [0, 0, 0], [0.4, 0.6, 1])) == 458800


def test_steps_for_frame_and_dt(orc):
    assert orc.steps_for_frame(0.0, 3, 24.0, 8) == 24
    assert orc.steps_for_frame(0.5, 3, 24.0, 8) == 120  # (0.5 + 3/24) * 24 * 8
    assert orc.steps_for_frame(0.0, -4, 24.0, 8) == 0
    assert orc.steps_for_frame(0.0, 0.5, 24.0, 1) == 1  # half away from zero, not to even
    assert orc.steps_for_frame(0.0, 2.5, 24.0, 1) == 3
    assert orc.dt_for(24.0, 5) == np.float32(1.0 / 120.0)
    assert orc.dt_for(30.0, 8) == np.float32(1.0 / 240.0)


# ---------------------------------------------------------------- binning

def test_bin_counting_sort_brute_force(orc):
    rng = np.random.default_rng(0)
    n, h = (7, 5, 6), np.float32(0.125)
    P = 3000
    x = rng.uniform(-0.1, 1.0, size=(P, 3)).astype(np.float32)  # some outside -> clamped
    sim = orc.Sim(n, h, [0, 0, 0], np.float32(0.01), 0.95, [0, -9.81, 0], x, np.zeros_like(x))
    cell_of, counts, offsets, ids = sim.bin()
    ci = np.floor((x - np.float32(0)) * (np.float32(1) / h)).astype(np.int64)
    ci = np.clip(ci, 0, np.array(n) - 1)
    lin = ci[:, 0] + n[0] * (ci[:, 1] + n[1] * ci[:, 2])
    np.testing.assert_array_equal(cell_of, lin)
    np.testing.assert_array_equal(counts, np.bincount(lin, minlength=np.prod(n)))
    np.testing.assert_array_equal(offsets, np.concatenate([[0], np.cumsum(counts)[:-1]]))
    np.testing.assert_array_equal(ids, np.argsort(lin, kind="stable"))


# ---------------------------------------------------------------- P2G

def _face_positions(n, h, origin):
    """World positions of every MAC face in reference storage order (u, v, w blocks)."""
    nx, ny, nz = n
    out = []
    for c in range(3):
        d = [nx, ny, nz]
        d[c] += 1
        k, j, i = np.meshgrid(np.arange(d[2]), np.arange(d[1]), np.arange(d[0]), indexing="ij")
        pos = np.stack([i, j, k], axis=-1).reshape(-1, 3).astype(np.float64) + 0.5
        pos[:, c] -= 0.5
        out.append(np.asarray(origin) + pos * h)
    return out


def test_p2g_partition_of_unity_and_first_moment(orc):
    """One particle: the fixed-point weights of each component sum to 1 and their first
    moment reproduces the particle position (trilinear interpolation is exact for linear
    functions), both to fixed-point precision; the normalized faces carry its velocity."""
    n, h = (10, 9, 8), 0.1
    rng = np.random.default_rng(1)
    faces = _face_positions(n, h, [0, 0, 0])
    for _ in range(20):
        x = rng.uniform([0.25, 0.25, 0.25], [0.75, 0.65, 0.55]).astype(np.float32)[None]
        v = rng.uniform(-2, 2, size=(1, 3)).astype(np.float32)
        sim = orc.Sim(n, h, [0, 0, 0], 0.01, 0.95, [0, 0, 0], x, v)
        sim.clear_grid()
        sim.p2g()
        sim.normalize_grid()
        off = 0
        for c in range(3):
            nf = len(faces[c])
            w = sim.wacc[off:off + nf].astype(np.float64)
            assert abs(w.sum() / 65536.0 - 1.0) < 8 / 65536.0
            m = (faces[c] * w[:, None]).sum(0) / w.sum()
            np.testing.assert_allclose(m, x[0], atol=2e-4 * h)
            nz = w > 0
            assert nz.sum() == 8
            # truncation bound: |trunc(w v S) / trunc(w S) - v| < (1 + |v|) / trunc(w S)
            err = np.abs(sim.vel_old[off:off + nf][nz].astype(np.float64) - v[0, c])
            assert np.all(err <= (1 + abs(float(v[0, c]))) / w[nz] + 1e-6)
            off += nf
        assert sim.count.sum() == 1


def test_p2g_skips_out_of_range_corners(orc):
    """A particle next to the low wall puts no weight outside the face grid (skipped, not
    clamped): the u weights still sum to 1 but the v/w weights lose the outside corners."""
    n, h = (6, 6, 6), 0.1
    x = np.float32([[0.02, 0.31, 0.33]])
    sim = orc.Sim(n, h, [0, 0, 0], 0.01, 0.95, [0, 0, 0], x, np.ones((1, 3), np.float32))
    sim.clear_grid()
    sim.p2g()
    nu = 7 * 6 * 6
    nv = 6 * 7 * 6
    wu = sim.wacc[:nu].sum() / 65536
    wv = sim.wacc[nu:nu + nv].sum() / 65536
    assert abs(wu - 1) < 1e-3
    # v faces: x-coordinate gx = 0.2 - 0.5 < 0 -> corner -1 skipped, weight 1 - 0.3 kept
    assert abs(wv - 0.7) < 1e-3


# ---------------------------------------------------------------- grid ops

def test_forces_closed_form(orc):
    n, h = (8, 8, 8), np.float32(0.125)
    dt = np.float32(1 / 120)
    x = np.float32([[0.5, 0.5, 0.5]])
    sim = orc.Sim(n, h, [0, 0, 0], dt, 0.95, [0, -9.81, 0], x, np.zeros((1, 3), np.float32))
    sim.clear_grid()
    sim.normalize_grid()
    sim.forces()
    nu, nv = 9 * 8 * 8, 8 * 9 * 8
    v = sim.vel[nu:nu + nv].reshape(8, 9, 8)  # [k][j][i]
    gdt = np.float32(0) + np.float32(-9.81) * dt
    # v face j sits between cells j-1 and j; both interior iff 2 <= j <= 6, i, k in 1..6
    inner = np.zeros_like(v, bool)
    inner[1:7, 2:7, 1:7] = True
    assert np.all(v[inner] == gdt)
    assert np.all(v[~inner] == 0)
    assert np.all(sim.vel[:nu] == 0) and np.all(sim.vel[nu + nv:] == 0)


def test_divergence_linear_field(orc):
    """u = a x, v = b y, w = c z on the faces -> div = a + b + c in every fluid cell."""
    n, h = (6, 7, 5), 0.2
    sim = orc.Sim(n, h, [0, 0, 0], 0.01, 0.95, [0, 0, 0], np.float32([[0.5, 0.5, 0.5]]), np.zeros((1, 3), np.float32))
    faces = _face_positions(n, h, [0, 0, 0])
    coef = [0.5, -1.25, 2.0]
    vel = np.concatenate([coef[c] * faces[c][:, c] for c in range(3)]).astype(np.float32)
    sim.vel[:] = vel
    sim.type[:] = 1
    sim.divergence()
    np.testing.assert_allclose(sim.div, sum(coef), rtol=1e-5, atol=1e-5)


# ---------------------------------------------------------------- pressure

def _slab(orc, res=16, layers=5, solver_iters=None):
    """Liquid filling the interior laterally, `layers` cells deep, at rest."""
    top = (layers + 1) / res - 1e-4
    f = dam_break(res, fill_max=(1.0, top, 1.0))
    sim = oracle_sim(orc, f)
    sim.grid_phase()
    n, h = sim.n, np.float32(1.0 / res)
    return sim, n, h


def _hydrostatic(sim, n, h, layers):
    p = np.zeros(sim.C)
    t = sim.type.reshape(n[2], n[1], n[0])
    assert np.all(t[1:-1, 1:layers + 1, 1:-1] == 1) and np.all(t[1:-1, layers + 1:-1, 1:-1] == 0)
    pj = np.zeros((n[2], n[1], n[0]))
    for j in range(1, layers + 1):
        pj[1:-1, j, 1:-1] = 9.81 * float(h) * (layers - j + 1)
    return pj.reshape(-1)


def test_hydrostatic_fp64_and_cg(orc):
    layers = 5
    sim, n, h = _slab(orc, layers=layers)
    exact = _hydrostatic(sim, n, h, layers)
    x, it, rr = sim.poisson_solve64(tol=1e-13)
    fl = sim.type == 1
    np.testing.assert_allclose(x[fl], exact[fl], rtol=1e-6)
    assert rr < 1e-12
    sim.pressure_cg32(60)
    np.testing.assert_allclose(sim.pressure[fl], exact[fl], rtol=5e-4)
    assert sim.poisson_relres64(sim.pressure) < 1e-4


def test_hydrostatic_rbgs_converges(orc):
    layers = 4
    sim, n, h = _slab(orc, res=12, layers=layers)
    exact = _hydrostatic(sim, n, h, layers)
    fl = sim.type == 1
    errs = []
    for _ in range(6):
        for _ in range(50):
            sim.pressure_rb(0)
            sim.pressure_rb(1)
        errs.append(np.abs(sim.pressure[fl] - exact[fl]).max())
    assert all(b <= a for a, b in zip(errs, errs[1:]))
    assert errs[-1] < 1e-3 * exact.max()


def test_hydrostatic_projection_zero_velocity(orc):
    """With the exact hydrostatic pressure the projected face velocities vanish."""
    layers = 5
    sim, n, h = _slab(orc, layers=layers)
    x, _, _ = sim.poisson_solve64(tol=1e-13)
    sim.pressure[:] = x.astype(np.float32)
    sim.project()
    nx, ny, nz = n
    nu, nv = (nx + 1) * ny * nz, nx * (ny + 1) * nz
    v = sim.vel[nu:nu + nv].reshape(nz, ny + 1, nx)
    # faces touching the liquid (v rows 1..layers+1) lose all velocity; air-air faces keep
    # g dt, as the reference leaves them (fluid.slang:192)
    assert np.abs(v[:, :layers + 2, :]).max() < 1e-5
    assert np.abs(sim.vel[:nu]).max() == 0 and np.abs(sim.vel[nu + nv:]).max() == 0


def test_poisson_matrix_vs_scipy(orc):
    """The FP64 oracle's matrix equals an independent sparse assembly of the reference
    stencil (diag = non-solid neighbours, -1 per fluid neighbour), and its solve agrees with
    scipy's direct solve."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    f = dam_break(12)
    sim = oracle_sim(orc, f)
    for _ in range(3):
        sim.step_rbgs(20)
    sim.grid_phase()
    n = sim.n
    t = sim.type.reshape(n[2], n[1], n[0])
    fluid = np.argwhere(t == 1)
    index = -np.ones(t.shape, np.int64)
    index[tuple(fluid.T)] = np.arange(len(fluid))
    rows, cols, vals = [], [], []
    for r, (k, j, i) in enumerate(fluid):
        diag = 0
        for dk, dj, di in [(0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)]:
            kk, jj, ii = k + dk, j + dj, i + di
            inside = 0 <= kk < n[2] and 0 <= jj < n[1] and 0 <= ii < n[0]
            tt = t[kk, jj, ii] if inside else 2
            if tt != 2:
                diag += 1
            if tt == 1:
                rows.append(r); cols.append(index[kk, jj, ii]); vals.append(-1.0)
        rows.append(r); cols.append(r); vals.append(float(diag))
    A = sp.csr_matrix((vals, (rows, cols)), shape=(len(fluid), len(fluid)))
    lin = fluid[:, 2] + n[0] * (fluid[:, 1] + n[1] * fluid[:, 0])
    rng = np.random.default_rng(3)
    xr = np.zeros(sim.C)
    xr[lin] = rng.standard_normal(len(fluid))
    np.testing.assert_allclose(sim.poisson_apply64(xr)[lin], A @ xr[lin], rtol=1e-13, atol=1e-12)
    b = sim.poisson_rhs64()[lin]
    xs = spl.spsolve(A.tocsc(), b)
    xo, it, rr = sim.poisson_solve64(tol=1e-13)
    np.testing.assert_allclose(xo[lin], xs, rtol=1e-8, atol=1e-9 * np.abs(xs).max())


def test_obstacle_cells_solid_and_faces_closed(orc):
    """Obstacle cells are SOLID after normalize (walls too; FLUID/AIR elsewhere from the
    count), every face touching one has zero velocity after forces, and the FP64 matrix over
    such a grid matches an independent assembly (diag = non-solid neighbours)."""
    import scipy.sparse as sp
    f = obstacle_scene(14)
    sim = oracle_sim(orc, f)
    sim.step_rbgs(10)
    sim.grid_phase()
    n = sim.n
    t = sim.type.reshape(n[2], n[1], n[0])
    ob = (sim.obst_id > 0).reshape(t.shape)
    assert ob.any() and np.all(t[ob] == 2)
    inner = np.zeros(t.shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    free = inner & ~ob
    np.testing.assert_array_equal(t[free], np.where(sim.count.reshape(t.shape)[free] > 0, 1, 0))
    nx, ny, nz = n
    nu, nv = (nx + 1) * ny * nz, nx * (ny + 1) * nz
    solid = np.pad(t == 2, 1, constant_values=True)  # cells outside the grid are SOLID
    u = sim.vel[:nu].reshape(nz, ny, nx + 1)
    vv = sim.vel[nu:nu + nv].reshape(nz, ny + 1, nx)
    w = sim.vel[nu + nv:].reshape(nz + 1, ny, nx)
    su = solid[1:-1, 1:-1, :-1] | solid[1:-1, 1:-1, 1:]
    sv = solid[1:-1, :-1, 1:-1] | solid[1:-1, 1:, 1:-1]
    sw = solid[:-1, 1:-1, 1:-1] | solid[1:, 1:-1, 1:-1]
    assert su.any() and not u[su].any() and not vv[sv].any() and not w[sw].any()
    assert u[~su].any() or vv[~sv].any()
    fluid = np.argwhere(t == 1)
    lin = fluid[:, 2] + nx * (fluid[:, 1] + ny * fluid[:, 0])
    pos = {tuple(c): q for q, c in enumerate(fluid)}
    rows, cols, vals = [], [], []
    for r, (k, j, i) in enumerate(fluid):
        diag = 0
        for dk, dj, di in [(0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)]:
            q = (k + dk, j + dj, i + di)
            if not solid[q[0] + 1, q[1] + 1, q[2] + 1]:
                diag += 1
            if q in pos:
                rows.append(r); cols.append(pos[q]); vals.append(-1.0)
        rows.append(r); cols.append(r); vals.append(float(diag))
    A = sp.csr_matrix((vals, (rows, cols)), shape=(len(fluid), len(fluid)))
    xr = np.zeros(sim.C)
    xr[lin] = np.random.default_rng(4).standard_normal(len(fluid))
    np.testing.assert_allclose(sim.poisson_apply64(xr)[lin], A @ xr[lin], rtol=1e-13, atol=1e-12)


def test_obstacle_push_invariants(orc):
    """Particles that end a G2P in an obstacle cell (zero grid velocity, so they stay put)
    move along the ray from that obstacle's centre to distance r + h; their new cell's centre
    is outside that sphere; an inward normal velocity is removed and the tangential part
    kept; an outward one is kept whole.  Particles elsewhere are bitwise those of the same
    G2P without obstacles."""
    n, h = (16, 16, 16), np.float32(1.0 / 16)
    centre, radius = np.float32([0.5, 0.45, 0.5]), np.float32(0.2)
    rng = np.random.default_rng(9)
    dirs = rng.standard_normal((400, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    x = (centre + dirs * rng.uniform(0.02, 0.3, (400, 1))).astype(np.float32)
    v = rng.uniform(-1, 1, (400, 3)).astype(np.float32)
    a = orc.Sim(n, h, [0, 0, 0], np.float32(0.01), np.float32(1.0), [0, 0, 0], x, v,
                obstacles=[(centre, radius)])
    b = orc.Sim(n, h, [0, 0, 0], np.float32(0.01), np.float32(1.0), [0, 0, 0], x, v)
    for s in (a, b):
        s.vel[:] = 0
        s.vel_old[:] = 0
        s.g2p_advect()
    cell = np.clip(np.floor(x * np.float32(16)).astype(int), 0, 15)
    inside = a.obst_id[cell[:, 0] + 16 * (cell[:, 1] + 16 * cell[:, 2])] > 0
    assert 20 < inside.sum() < 380
    np.testing.assert_array_equal(a.x[~inside], b.x[~inside])
    np.testing.assert_array_equal(a.v[~inside], b.v[~inside])
    d0 = (x[inside] - centre).astype(np.float64)
    d1 = (a.x[inside] - centre).astype(np.float64)
    nr = d0 / np.linalg.norm(d0, axis=1, keepdims=True)
    np.testing.assert_allclose(np.linalg.norm(d1, axis=1), radius + h, rtol=1e-6)
    np.testing.assert_allclose(d1 / np.linalg.norm(d1, axis=1, keepdims=True), nr, atol=1e-6)
    c1 = np.clip(np.floor(a.x[inside] * np.float32(16)).astype(int), 0, 15)
    centres = (c1 + 0.5) / 16.0
    assert np.all(np.linalg.norm(centres - centre, axis=1) > radius)
    v0 = v[inside].astype(np.float64)
    v1 = a.v[inside].astype(np.float64)
    vn0 = np.sum(v0 * nr, axis=1)
    np.testing.assert_allclose(np.sum(v1 * nr, axis=1), np.maximum(vn0, 0), atol=1e-6)
    tang0 = v0 - vn0[:, None] * nr
    np.testing.assert_allclose(v1 - np.sum(v1 * nr, axis=1)[:, None] * nr, tang0, atol=1e-6)
    out = vn0 >= 0
    assert out.any() and (~out).any()
    np.testing.assert_array_equal(a.v[inside][out], v[inside][out])


def test_obstacle_scene_exercises_the_push(orc):
    """The obstacle parity scene (4 reference steps at res 24) moves particles out of
    obstacles: after each step some sit exactly at distance r + h from an obstacle centre
    (pushed in that step), and (almost) none ends a step in an obstacle cell (only a push
    into an overlapping obstacle can leave one there)."""
    f = obstacle_scene(24)
    sim = oracle_sim(orc, f, seed=7)
    pushed, stuck = 0, 0
    for _ in range(4):
        sim.step_rbgs(40)
        x = sim.x.astype(np.float64)
        for o in f["obstacles"]:
            d = np.linalg.norm(x - np.asarray(o["center"]), axis=1)
            pushed += int(np.sum(np.abs(d - (o["radius"] + float(sim.params.h))) < 2e-6))
        cell = np.clip(np.floor(sim.x * np.float32(24)).astype(int), 0, 23)
        stuck += int((sim.obst_id[cell[:, 0] + 24 * (cell[:, 1] + 24 * cell[:, 2])] > 0).sum())
    assert pushed > 20 and stuck <= 2


def test_rbgs_half_sweep_touches_one_colour(orc):
    sim = oracle_sim(orc, dam_break(12))
    sim.grid_phase()
    before = sim.pressure.copy()
    sim.pressure_rb(1)
    n = sim.n
    k, j, i = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    parity = ((i + j + k) % 2).reshape(-1)
    changed = sim.pressure != before
    assert changed.any() and not changed[parity == 0].any()


# ---------------------------------------------------------------- G2P

def test_g2p_uniform_field(orc):
    """Uniform grid field c, vel_old = c: PIC = c, FLIP delta = 0, so v' = c(1-t) + v t and
    x' = x + dt c (no wall contact)."""
    n, h, dt, t = (10, 10, 10), np.float32(0.1), np.float32(0.01), np.float32(0.7)
    rng = np.random.default_rng(2)
    x = rng.uniform(0.3, 0.7, size=(50, 3)).astype(np.float32)
    v = rng.uniform(-1, 1, size=(50, 3)).astype(np.float32)
    sim = orc.Sim(n, h, [0, 0, 0], dt, t, [0, 0, 0], x, v)
    c = np.float32([0.3, -0.2, 0.1])
    nu, nv = 11 * 100, 110 * 10
    sim.vel[:nu] = c[0]; sim.vel[nu:nu + nv] = c[1]; sim.vel[nu + nv:] = c[2]
    sim.vel_old[:] = sim.vel
    sim.g2p_advect()
    np.testing.assert_allclose(sim.v, c * (1 - t) + v * t, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(sim.x, x + dt * c, rtol=0, atol=1e-6)


def test_g2p_linear_field_rk2_midpoint(orc):
    """A linear grid field u = a (x - x0), v = b (y - y0), w = 0 (vel_old = vel, flip = 0), which
    trilinear sampling reproduces away from the walls: v' = PIC = the field at x, and the RK2
    midpoint step gives x' - x0 = (x - x0)(1 + a dt + (a dt)^2 / 2) exactly (y likewise with b).
    An Euler step (sampling at x instead of the midpoint) misses the (a dt)^2 / 2 term, which
    here is ~700 f32 ulps of the position: the test tells the two apart."""
    n, h, dt = (10, 10, 10), np.float32(0.1), np.float32(0.01)
    a, b, x0, y0 = 3.0, -2.0, 0.45, 0.55
    rng = np.random.default_rng(5)
    x = rng.uniform(0.3, 0.7, size=(200, 3)).astype(np.float32)
    sim = orc.Sim(n, h, [0, 0, 0], dt, 0.0, [0, 0, 0], x, np.zeros_like(x))
    fu, fv, fw = _face_positions(n, float(h), [0, 0, 0])
    field = np.concatenate([a * (fu[:, 0] - x0), b * (fv[:, 1] - y0), np.zeros(len(fw))]).astype(np.float32)
    sim.vel[:] = field
    sim.vel_old[:] = field
    sim.g2p_advect()
    X = x.astype(np.float64)
    d = float(dt)
    np.testing.assert_allclose(sim.vx, a * (X[:, 0] - x0), atol=2e-6)
    np.testing.assert_allclose(sim.vy, b * (X[:, 1] - y0), atol=2e-6)
    assert np.abs(sim.vz).max() < 1e-6
    rk2x = x0 + (X[:, 0] - x0) * (1 + a * d + (a * d) ** 2 / 2)
    rk2y = y0 + (X[:, 1] - y0) * (1 + b * d + (b * d) ** 2 / 2)
    np.testing.assert_allclose(sim.px, rk2x, rtol=0, atol=1e-6)
    np.testing.assert_allclose(sim.py, rk2y, rtol=0, atol=1e-6)
    np.testing.assert_array_equal(sim.pz, x[:, 2])
    euler = x0 + (X[:, 0] - x0) * (1 + a * d)
    far = np.abs(X[:, 0] - x0) > 0.05
    assert np.all(np.abs(sim.px[far] - euler[far]) > 2e-5)


def test_g2p_clamps_the_stencil_where_p2g_skips(orc):
    """Near the low x wall the v- and w-face stencils reach face column -1: sample_face clamps
    that corner onto column 0 (fluid.slang:75), where p2g skips it (fluid.slang:105).  A
    constant field c on the v faces is therefore sampled as c there (weights 0.3 + 0.7 on column
    0, to f32 rounding), not 0.7 c; the same particle's P2G keeps only the 0.7 in-range weight."""
    n, h, dt = (6, 6, 6), np.float32(0.1), np.float32(0.001)
    x = np.float32([[0.02, 0.35, 0.33]])
    c = np.float32(0.8)
    sim = orc.Sim(n, h, [0, 0, 0], dt, 0.0, [0, 0, 0], x, np.zeros((1, 3), np.float32))
    nu, nv = 7 * 6 * 6, 6 * 7 * 6
    sim.vel[:] = 0
    sim.vel[nu:nu + nv] = c
    sim.vel_old[:] = sim.vel
    sim.g2p_advect()
    assert abs(float(sim.vy[0]) - float(c)) < 1e-6  # f32 rounding of the weight sums; skipping gives 0.56
    assert sim.vx[0] == 0 and sim.vz[0] == 0
    sim.set_particles(x, np.ones((1, 3), np.float32))
    sim.clear_grid()
    sim.p2g()
    assert abs(sim.wacc[nu:nu + nv].sum() / 65536 - 0.7) < 1e-3


def test_g2p_wall_clamp(orc):
    n, h, dt = (8, 8, 8), np.float32(0.125), np.float32(0.1)
    x = np.float32([[0.2, 0.5, 0.5], [0.85, 0.5, 0.5]])
    v = np.float32([[-3, 0, 0], [3, 0, 0]])
    sim = orc.Sim(n, h, [0, 0, 0], dt, 1.0, [0, 0, 0], x, v)
    sim.vel[:9 * 64] = np.float32(-20.0)  # strong flow toward -x everywhere
    sim.vel_old[:] = 0
    sim.g2p_advect()
    lo = np.float32(0) + h * np.float32(1.001)
    assert sim.px[0] == lo and sim.vx[0] == 0.0  # pushed into the low wall: clamp + v = max(v, 0)
    assert sim.px[1] >= lo
    hi = np.float32(0) + h * (np.float32(8) - np.float32(1.001))
    assert np.all(sim.x <= hi) and np.all(sim.x >= lo)


def test_free_fall_flip1(orc):
    """A blob away from the walls with flip = 1 falls freely: after n steps every particle
    has v_y = -n g dt (1e-3 relative: fixed-point P2G truncation at the blob surface)
    and v_x = v_z = 0, and the pressure stays ~0."""
    f = {"name": "blob", "bounds": [[0, 0, 0], [1, 1, 1]], "resolution": 24,
         "fill": [{"min": [0.35, 0.6, 0.35], "max": [0.65, 0.8, 0.65]}],
         "flip": 1.0, "substeps": 8, "particles_per_cell": 4}
    sim = oracle_sim(orc, f)
    dt = orc.dt_for(FPS, 8)
    for s in range(1, 5):
        sim.step_rbgs(80)
        np.testing.assert_allclose(sim.vy, -s * 9.81 * float(dt), rtol=1e-3)
        vfall = s * 9.81 * float(dt)
        assert np.abs(sim.vx).max() < 1e-3 * vfall and np.abs(sim.vz).max() < 1e-3 * vfall
        # a resting blob of this height would carry |g| * 0.2 ~ 2; free fall keeps ~0
        assert np.abs(sim.pressure).max() < 2e-3 * 9.81 * 0.2


def test_cfl_substeps_closed_form(orc):
    """CFL cap (an extension): k = ceil(|v|max dt / (cfl h)), 1 at or below one cell, at
    most 64; |v|max is the largest particle speed."""
    x = np.full((3, 3), 0.5, np.float32)
    v = np.array([[3, 4, 0], [1, 1, 1], [0, -2, 0]], np.float32)  # |v|max = 5
    s = orc.Sim((32, 32, 32), np.float32(1 / 32), [0, 0, 0], np.float32(1 / 120), 0.95, [0, -9.81, 0], x, v)
    assert s.cfl_substeps(1.0) == 2  # 5 * (1/120) * 32 = 1.33 cells
    assert s.cfl_substeps(0.5) == 3  # 2.67
    assert s.cfl_substeps(2.0) == 1  # 0.67
    assert s.cfl_substeps(1e-3) == 64
    s.vx[:] = 0
    s.vy[:] = 0
    s.vz[:] = 0
    assert s.cfl_substeps(1e-3) == 1


def test_cfl_capped_free_fall(orc):
    """A blob thrown down fast enough to need 3 substeps: the velocity change is -g dt in
    total (k steps of dt / k); inside the blob the displacement is the free-fall one, and at
    the blob's leading face, where one long RK2 step samples the air below, the substeps
    stay closer to it than the uncapped step; a step slow enough for k = 1 is bitwise the
    reference step."""
    f = {"name": "blob", "bounds": [[0, 0, 0], [1, 1, 1]], "resolution": 24,
         "fill": [{"min": [0.35, 0.5, 0.35], "max": [0.65, 0.7, 0.65], "velocity": [0, -6.0, 0]}],
         "flip": 1.0, "substeps": 5, "particles_per_cell": 4}
    a, b = oracle_sim(orc, f), oracle_sim(orc, f)
    dt = float(orc.dt_for(FPS, 5))
    y0 = a.py.astype(np.float64).copy()
    assert a.cfl_substeps(1.0) == 2 and a.cfl_substeps(0.45) == 3  # 6 m/s * (1/120) s * 24 / m = 1.2 cells
    k = a.step_rbgs_capped(80, 0.45)
    b.step_rbgs(80)
    assert k == 3
    np.testing.assert_allclose(a.vy, -6.0 - 9.81 * dt, rtol=2e-3)
    # inside the blob each (sub)step moves by its length times the post-gravity velocity:
    # k substeps give -6 dt - g dt^2 (k + 1) / (2 k), one step -6 dt - g dt^2
    da = a.py - y0 - (-6.0 * dt - 9.81 * dt * dt * 4 / 6)
    db = b.py - y0 - (-6.0 * dt - 9.81 * dt * dt)
    assert np.median(np.abs(da)) < 1e-5 and np.median(np.abs(db)) < 1e-5  # f32, fixed-point P2G
    assert np.abs(da).max() < 0.5 * np.abs(db).max()
    c, d = oracle_sim(orc, f), oracle_sim(orc, f)
    assert c.step_rbgs_capped(80, 100.0) == 1
    d.step_rbgs(80)
    np.testing.assert_array_equal(c.x, d.x)
    np.testing.assert_array_equal(c.v, d.v)


def test_whole_step_hydrostatic_rest(orc):
    """A resting slab stays at rest through full steps (pressure balances gravity)."""
    f = dam_break(16, fill_max=(1.0, 0.3, 1.0))
    sim = oracle_sim(orc, f)
    for _ in range(2):
        sim.step_cg(200)
    assert np.abs(sim.v).max() < 2e-3
