"""Pins the oracle's surface nets (oracle/flip_oracle.c, restating shaders/fluid.slang:222-355
and crates/pxl-render/src/fluid.rs:151-157, 308-330) to facts that do not come from it:

  parameters  closed forms of sx, sh, radius, iso (the poly6-style kernel volume
              4 pi * 16/315 = 0.638 in node units);
  splat       hand-derived node values of one particle on a node (q = 1, 3/4, 1/2, 1/4 at
              squared distances 0..3: 4096, 1728, 512, 64) and the lattice-sum mass of random
              particles (its mean over a uniform offset is the kernel's integral);
  phi         constant density gives density / 4096 - iso inside and -1 on the boundary;
  mesh        a linear field (vertices on the plane, normals -grad / |grad|), a sphere
              (a closed, consistently wound 2-manifold with Euler characteristic 2, outward
              normals, the enclosed volume) and the overflow rules.
"""
import numpy as np
import pytest


def _params(orc, n=(16, 16, 16), h=1.0 / 16, refine=2, radius=2.0, ppc=8, max_tris=2_000_000):
    return orc.surface_params(n, np.float32(h), [0, 0, 0], refine, radius, ppc, max_tris)


def test_surface_params_closed_forms(orc):
    sp = _params(orc, n=(64, 32, 48), h=1.0 / 64)
    assert (sp.sx, sp.sy, sp.sz) == (128, 64, 96)
    assert sp.sh == np.float32(1.0 / 128) and sp.radius == np.float32(2.0 / 128)
    assert sp.iso == pytest.approx(0.5 * 8 * 4 * np.pi * 16 / 315, rel=2e-3)
    sp3 = _params(orc, n=(20, 20, 20), h=0.05, refine=3, radius=2.0, ppc=5)
    assert sp3.sx == 60 and sp3.iso == pytest.approx(0.5 * 5 * 0.638 * (2 / 3) ** 3, rel=1e-6)
    assert sp3.max_verts == sp3.max_tris == 2_000_000


def test_splat_one_particle_on_a_node(orc):
    sp = _params(orc)
    node = np.array([5, 7, 9])
    x = (node.astype(np.float32) * np.float32(sp.sh))[None, :]
    d = orc.surface_splat(sp, x).reshape(sp.sz + 1, sp.sy + 1, sp.sx + 1)
    nz = np.argwhere(d)
    r2 = ((nz[:, ::-1] - node) ** 2).sum(axis=1)
    expect = {0: 4096, 1: 1728, 2: 512, 3: 64}
    assert sorted(set(r2.tolist())) == [0, 1, 2, 3] and len(nz) == 27
    for (k, j, i), q in zip(nz, r2):
        assert d[k, j, i] == expect[int(q)]
    assert d.sum() == 4096 + 6 * 1728 + 12 * 512 + 8 * 64


def test_splat_mass_of_random_particles(orc):
    sp = _params(orc, n=(24, 24, 24), h=1.0 / 24)
    rng = np.random.default_rng(5)
    x = rng.uniform(0.25, 0.75, (4000, 3)).astype(np.float32)
    d = orc.surface_splat(sp, x)
    r = 2.0
    integral = 4 * np.pi * 16 / 315 * r ** 3 * 4096  # per particle, node units
    # int() truncates: each of the ~4/3 pi r^3 nodes a particle touches loses half a unit
    expect = len(x) * (integral - 0.5 * 4 / 3 * np.pi * r ** 3)
    assert abs(d.sum() / expect - 1) < 3e-3


def test_phi_of_constant_density(orc):
    sp = _params(orc, n=(8, 6, 10), h=0.125)
    dens = np.full(sp.nodes, 12288, np.int32)
    phi = orc.surface_phi(sp, dens).reshape(sp.sz + 1, sp.sy + 1, sp.sx + 1)
    inner = phi[1:-1, 1:-1, 1:-1]
    assert np.all(inner == np.float32(3.0) - np.float32(sp.iso))
    edge = np.ones(phi.shape, bool)
    edge[1:-1, 1:-1, 1:-1] = False
    assert np.all(phi[edge] == -1.0)


def _node_grid(sp):
    k, j, i = np.meshgrid(np.arange(sp.sz + 1), np.arange(sp.sy + 1), np.arange(sp.sx + 1), indexing="ij")
    return np.stack([i, j, k], axis=-1).reshape(-1, 3).astype(np.float64)


def test_mesh_of_a_linear_field(orc):
    """Surface nets of phi = b - a.p: every vertex lies on the plane a.p = b and every
    normal is a / |a| (outward = toward decreasing phi)."""
    sp = _params(orc, n=(6, 6, 6), h=1.0 / 6)
    a = np.array([0.3, 1.0, -0.45])
    b = 6.3
    p = _node_grid(sp)
    phi = (b - p @ a).astype(np.float32)
    cv, verts, tris, (nv, nt) = orc.surface_mesh(sp, phi)
    assert nv == len(verts) > 50 and nt == len(tris) > 50
    local = verts[:, :3].astype(np.float64) / float(sp.sh)
    np.testing.assert_allclose(local @ a, b, atol=2e-5)
    np.testing.assert_allclose(verts[:, 4:7], np.tile(a / np.linalg.norm(a), (nv, 1)), atol=2e-6)
    assert not verts[:, 3].any() and not verts[:, 7].any()
    # every triangle's geometric normal agrees with the outward direction
    v = verts[:, :3].astype(np.float64)
    gn = np.cross(v[tris[:, 1]] - v[tris[:, 0]], v[tris[:, 2]] - v[tris[:, 0]])
    assert np.all(gn @ a > 0)


def _sphere_mesh(orc, R=5.3, c=(8.2, 7.9, 8.4), n=8):
    sp = _params(orc, n=(n, n, n), h=1.0 / n)
    p = _node_grid(sp)
    phi = (R - np.linalg.norm(p - np.asarray(c), axis=1)).astype(np.float32)
    return sp, orc.surface_mesh(sp, phi), np.asarray(c) * float(sp.sh), R * float(sp.sh)


def test_mesh_of_a_sphere_is_a_closed_oriented_manifold(orc):
    sp, (cv, verts, tris, (nv, nt)), centre, R = _sphere_mesh(orc)
    assert nv == len(verts) and nt == len(tris) and nt > 100
    # each directed edge once, each undirected edge in exactly two triangles
    directed = np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]])
    assert len({tuple(e) for e in directed.tolist()}) == len(directed)
    und = np.sort(directed, axis=1)
    _, counts = np.unique(und, axis=0, return_counts=True)
    assert np.all(counts == 2)
    E = len(counts)
    used = np.unique(tris)
    assert len(used) == nv and nv - E + nt == 2
    v = verts[:, :3].astype(np.float64)
    dist = np.linalg.norm(v - centre, axis=1)
    assert np.all(np.abs(dist - R) < 0.5 * float(sp.sh))
    assert np.all(np.sum(verts[:, 4:7] * (v - centre), axis=1) > 0)
    vol = np.einsum("ij,ij->i", v[tris[:, 0]], np.cross(v[tris[:, 1]], v[tris[:, 2]])).sum() / 6
    # positive: outward winding; a little below the ball's volume: each vertex is a mean of
    # points on the sphere, so the polyhedron lies inside it
    exact = 4 / 3 * np.pi * R ** 3
    assert 0.92 * exact < vol < exact


def test_mesh_overflow_rules(orc):
    """Vertices past max_verts are not stored (their cells hold none, so quads needing them
    are dropped); the counters keep counting; triangles past max_tris are not written."""
    sp, (_, verts, tris, (nv, nt)), _, _ = _sphere_mesh(orc)
    sp.max_verts = nv // 2
    sp.max_tris = 40
    p = _node_grid(sp)
    phi = (5.3 - np.linalg.norm(p - np.array([8.2, 7.9, 8.4]), axis=1)).astype(np.float32)
    cv, v2, t2, (nv2, nt2) = orc.surface_mesh(sp, phi)
    assert nv2 == nv and len(v2) == nv // 2
    np.testing.assert_array_equal(v2, verts[:nv // 2])
    assert nt2 < nt and len(t2) == 40 and t2.max() < nv // 2
    assert np.all(cv[cv != 0xFFFFFFFF] < nv // 2)
