"""GPU surface nets (shaders/fluid.slang:222-355) against the oracle's restatement on the
same particles: density (integer sums), phi, the cell -> vertex map, every vertex (position
and normal) and every triangle bit for bit -- the kernels number vertices in surface-cell
order and triangles in (node, axis) order, the order of the oracle's loops -- plus the
reference's counters and clipping rules."""
import numpy as np
import pytest

from tests.scenes import EXAMPLE_DAM_BREAK, FPS, dam_break, obstacle_scene, sphere_drop

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pxl():
    import pxlengine_amd
    return pxlengine_amd


def _oracle_surface(orc, gpu, fluid):
    info = gpu.info
    sp = orc.surface_params(info["n"], info["h"], fluid["bounds"][0], fluid.get("surface_refine", 2),
                            fluid.get("particle_radius", 2.0), fluid.get("particles_per_cell", 8),
                            fluid.get("max_triangles", 2_000_000))
    p = gpu.read_particles()
    dens = orc.surface_splat(sp, p[:, 0:3])
    phi = orc.surface_phi(sp, dens)
    cv, verts, tris, ctr = orc.surface_mesh(sp, phi)
    return sp, dens, phi, cv, verts, tris, ctr


def _check(orc, gpu, fluid, what):
    print(f"surface {what}", flush=True)
    verts, tris, info = gpu.surface()
    sp, dens, phi, cv, overts, otris, ctr = _oracle_surface(orc, gpu, fluid)
    assert (info["sx"], info["sy"], info["sz"]) == (sp.sx, sp.sy, sp.sz)
    assert info["iso"] == sp.iso and info["sh"] == sp.sh and info["radius"] == sp.radius
    np.testing.assert_array_equal(gpu.read_field("density"), dens, err_msg=f"density {what}")
    np.testing.assert_array_equal(gpu.read_field("phi"), phi, err_msg=f"phi {what}")
    np.testing.assert_array_equal(gpu.read_field("cell_vert"), cv, err_msg=f"cell_vert {what}")
    assert (info["vertices_total"], info["triangles_total"]) == ctr
    np.testing.assert_array_equal(verts, overts, err_msg=f"vertices {what}")
    np.testing.assert_array_equal(tris, otris, err_msg=f"triangles {what}")
    assert info["overflow"] == int(ctr[0] > sp.max_verts or ctr[1] > sp.max_tris)
    return verts, tris, info


@pytest.mark.parametrize("fluid,steps", [
    (dam_break(24), 3),
    (dict(dam_break(20), surface_refine=3, particle_radius=2.5), 2),
    (dict(dam_break(32, ppc=4), surface_refine=1, particle_radius=1.3), 2),
    (EXAMPLE_DAM_BREAK, 2),
    (sphere_drop(28), 3),
    (obstacle_scene(24), 3),
    (dam_break(64, ppc=1), 1),
], ids=["dam24", "refine3", "refine1", "example", "sphere28", "obstacles24", "sparse64"])
def test_surface_bitwise(pxl, orc, fluid, steps):
    """(sparse64: one particle per cell, so a 2048-particle splat chunk spans more nodes than
    its LDS box holds and takes the global-atomic path.)"""
    gpu = pxl.FluidState(fluid, FPS, seed=0)
    _check(orc, gpu, fluid, "at seeding")
    gpu.step(steps)
    verts, tris, info = _check(orc, gpu, fluid, f"after {steps} steps")
    assert info["vertices"] > 100 and info["triangles"] > 100 and not info["overflow"]
    assert tris.max() < info["vertices"]
    gpu.close()


def test_surface_overflow_clipping(pxl, orc):
    """max_triangles below the mesh's size: vertices past it are not stored, quads needing
    them are dropped, triangles past it are not written, overflow is reported -- as the
    oracle does it."""
    fluid = dict(dam_break(24), max_triangles=1500)
    gpu = pxl.FluidState(fluid, FPS, seed=0)
    gpu.step(2)
    verts, tris, info = _check(orc, gpu, fluid, "clipped")
    assert info["overflow"] == 1 and info["vertices"] == 1500 and info["vertices_total"] > 1500
    assert len(tris) <= 1500 and len(tris) % 2 == 0
    gpu.close()


def test_surface_into_device_buffers(pxl):
    """The device entry point writes the same mesh at vbase / ibase of caller buffers and
    leaves the rest of them untouched."""
    fluid = dam_break(24)
    gpu = pxl.FluidState(fluid, FPS, seed=0)
    gpu.step(2)
    verts, tris, info = gpu.surface()
    cap = gpu.reservation()[0]
    vbase, ibase = 1000, 777
    vhost = np.full((vbase + cap, 8), -7.0, np.float32)
    ihost = np.full(ibase + 3 * cap, 0xABCDEF, np.uint32)
    vdev, idev = pxl.DeviceBuffer(vhost.nbytes), pxl.DeviceBuffer(ihost.nbytes)
    try:
        vdev.upload(vhost)
        idev.upload(ihost)
        info2 = gpu.surface_device(vdev.ptr, vbase, idev.ptr, ibase)
        v = vdev.download(np.empty_like(vhost))
        ib = idev.download(np.empty_like(ihost))
    finally:
        vdev.free()
        idev.free()
    assert info2 == info
    np.testing.assert_array_equal(v[vbase:vbase + info["vertices"]], verts)
    np.testing.assert_array_equal(ib[ibase:ibase + 3 * info["triangles"]].reshape(-1, 3), tris)
    assert np.all(v[:vbase] == -7.0) and np.all(v[vbase + info["vertices"]:] == -7.0)
    assert np.all(ib[:ibase] == 0xABCDEF) and np.all(ib[ibase + 3 * info["triangles"]:] == 0xABCDEF)
    gpu.close()


def test_surface_full_size_dam(pxl, orc):
    """The 128^3 dam break (3.83M particles) after 5 steps, surface grid 256^3 (refine 2)."""
    fluid = dam_break(128)
    gpu = pxl.FluidState(fluid, FPS, seed=0, solver="cg", cg_iterations=80)
    gpu.step(5)
    verts, tris, info = _check(orc, gpu, fluid, "128^3")
    assert info["vertices"] > 10_000
    gpu.close()


class _Exporter:
    """A device allocation made exportable as a POSIX file descriptor through HIP's virtual
    memory API -- standing in for a Vulkan buffer exported with VK_KHR_external_memory_fd --
    mapped here too so the test can read what the importer wrote."""

    def __init__(self, nbytes):
        import ctypes as C
        import os
        lib = C.CDLL("/opt/rocm/lib/libamdhip64.so" if os.path.exists("/opt/rocm/lib/libamdhip64.so")
                     else "libamdhip64.so")

        class Loc(C.Structure):
            _fields_ = [("type", C.c_int), ("id", C.c_int)]

        class Flags(C.Structure):
            _fields_ = [("compressionType", C.c_ubyte), ("gpuDirectRDMACapable", C.c_ubyte), ("usage", C.c_ushort)]

        class Prop(C.Structure):
            _fields_ = [("type", C.c_int), ("requestedHandleType", C.c_int), ("location", Loc),
                        ("win32HandleMetaData", C.c_void_p), ("allocFlags", Flags)]

        class Access(C.Structure):
            _fields_ = [("location", Loc), ("flags", C.c_int)]

        self.C, self.lib = C, lib
        supported = C.c_int(0)
        # hipDeviceAttributeVirtualMemoryManagementSupported
        lib.hipDeviceGetAttribute.argtypes = [C.POINTER(C.c_int), C.c_int, C.c_int]
        prop = Prop(type=1, requestedHandleType=1, location=Loc(type=1, id=0))  # pinned, POSIX fd, device 0
        gran = C.c_size_t(0)
        lib.hipMemGetAllocationGranularity.argtypes = [C.POINTER(C.c_size_t), C.POINTER(Prop), C.c_int]
        if lib.hipMemGetAllocationGranularity(C.byref(gran), C.byref(prop), 0) != 0 or gran.value == 0:
            pytest.skip("HIP virtual memory management is not available here")
        self.size = (nbytes + gran.value - 1) // gran.value * gran.value
        self.handle = C.c_uint64(0)
        lib.hipMemCreate.argtypes = [C.POINTER(C.c_uint64), C.c_size_t, C.POINTER(Prop), C.c_ulonglong]
        if lib.hipMemCreate(C.byref(self.handle), self.size, C.byref(prop), 0) != 0:
            pytest.skip("hipMemCreate with an exportable handle is not available here")
        fd = C.c_int(-1)
        lib.hipMemExportToShareableHandle.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_ulonglong]
        if lib.hipMemExportToShareableHandle(C.byref(fd), self.handle, 1, 0) != 0:
            pytest.skip("exporting a HIP allocation as a file descriptor is not available here")
        self.fd = fd.value
        self.va = C.c_void_p()
        lib.hipMemAddressReserve.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_size_t, C.c_void_p, C.c_ulonglong]
        lib.hipMemMap.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_uint64, C.c_ulonglong]
        lib.hipMemSetAccess.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(Access), C.c_size_t]
        assert lib.hipMemAddressReserve(C.byref(self.va), self.size, 0, None, 0) == 0
        assert lib.hipMemMap(self.va, self.size, 0, self.handle, 0) == 0
        acc = Access(location=Loc(type=1, id=0), flags=3)
        assert lib.hipMemSetAccess(self.va, self.size, C.byref(acc), 1) == 0
        lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]

    def write(self, host):
        assert self.lib.hipMemcpy(self.va, host.ctypes.data, host.nbytes, 1) == 0

    def read(self, out):
        assert self.lib.hipMemcpy(out.ctypes.data, self.va, out.nbytes, 2) == 0
        return out

    def close(self):
        C, lib = self.C, self.lib
        lib.hipMemUnmap.argtypes = [C.c_void_p, C.c_size_t]
        lib.hipMemAddressFree.argtypes = [C.c_void_p, C.c_size_t]
        lib.hipMemRelease.argtypes = [C.c_uint64]
        lib.hipMemUnmap(self.va, self.size)
        lib.hipMemAddressFree(self.va, self.size)
        lib.hipMemRelease(self.handle)


def test_surface_into_imported_buffers(pxl):
    """Interop: the vertex and index buffers are allocations exported as file descriptors
    (as Vulkan exports the renderer's buffers); the library imports them, meshes into them at
    vbase / ibase, and the exporter's own mapping sees the mesh."""
    fluid = dam_break(24)
    gpu = pxl.FluidState(fluid, FPS, seed=0)
    gpu.step(2)
    verts, tris, info = gpu.surface()
    cap = gpu.reservation()[0]
    vbase, ibase = 64, 96
    vex, iex = _Exporter(32 * (vbase + cap)), _Exporter(4 * (ibase + 3 * cap))
    vimp = iimp = None
    try:
        vhost = np.full((vbase + cap, 8), -7.0, np.float32)
        vex.write(vhost)
        vimp = pxl.ImportedBuffer(vex.fd, vex.size)
        iimp = pxl.ImportedBuffer(iex.fd, iex.size)
        info2 = gpu.surface_device(vimp.ptr, vbase, iimp.ptr, ibase)
        assert info2 == info
        v = vex.read(np.empty((vbase + cap, 8), np.float32))
        ib = iex.read(np.empty(ibase + 3 * cap, np.uint32))
        np.testing.assert_array_equal(v[vbase:vbase + info["vertices"]], verts)
        np.testing.assert_array_equal(ib[ibase:ibase + 3 * info["triangles"]].reshape(-1, 3), tris)
        assert np.all(v[:vbase] == -7.0)
    finally:
        for x in (vimp, iimp):
            if x is not None:
                x.release()
        vex.close()
        iex.close()
    gpu.close()


@pytest.mark.parametrize("dedicated", [False, True], ids=["suballocated", "dedicated-flag"])
def test_surface_into_suballocated_buffers(pxl, dedicated):
    """Interop with a suballocator (the renderer's buffers are ranges of one VkDeviceMemory,
    crates/pxl-gpu/src/lib.rs:186-213): one exported allocation is imported once and the vertex
    and index buffers are mapped from it at non-zero memory offsets; meshing into them at
    vbase / ibase writes exactly the mesh at offset + base, and the rest of the block (before,
    between and after the buffers) is untouched, as the exporter's own mapping shows."""
    import os
    fluid = dam_break(24)
    gpu = pxl.FluidState(fluid, FPS, seed=0)
    gpu.step(2)
    verts, tris, info = gpu.surface()
    cap = gpu.reservation()[0]
    vbase, ibase = 40, 72
    vbytes, ibytes = 32 * (vbase + cap), 4 * (ibase + 3 * cap)
    voff = 65536
    ioff = (voff + vbytes + 65535) // 65536 * 65536 + 256
    ex = _Exporter(ioff + ibytes + 4096)
    mem = None
    try:
        fill = np.full(ex.size // 4, 0x7F7F7F7F, np.uint32)
        ex.write(fill)
        mem = pxl.ImportedMemory(os.dup(ex.fd), ex.size, dedicated=dedicated)
        vptr, iptr = mem.map(voff, vbytes), mem.map(ioff, ibytes)
        with pytest.raises(pxl.FluidError):
            mem.map(ex.size - 16, 32)  # past the end of the allocation
        info2 = gpu.surface_device(vptr, vbase, iptr, ibase)
        assert info2 == info
        got = ex.read(np.empty(ex.size // 4, np.uint32))
        vw = got[voff // 4 + 8 * vbase: voff // 4 + 8 * (vbase + info["vertices"])].view(np.float32).reshape(-1, 8)
        iw = got[ioff // 4 + ibase: ioff // 4 + ibase + 3 * info["triangles"]].reshape(-1, 3)
        np.testing.assert_array_equal(vw, verts)
        np.testing.assert_array_equal(iw, tris)
        touched = np.zeros(len(got), bool)
        touched[voff // 4 + 8 * vbase: voff // 4 + 8 * (vbase + info["vertices"])] = True
        touched[ioff // 4 + ibase: ioff // 4 + ibase + 3 * info["triangles"]] = True
        assert np.all(got[~touched] == 0x7F7F7F7F)
    finally:
        if mem is not None:
            mem.release()
        ex.close()
    gpu.close()
