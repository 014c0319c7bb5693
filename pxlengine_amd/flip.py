"""ctypes binding of ``libpxlflip.so`` (the C ABI in ``include/pxlflip.h``).

``FluidState`` mirrors the reference's per-fluid driver, ``FluidState`` in
crates/pxl-render/src/fluid.rs:74-306: it is created from one scene ``fluids[]`` block plus
``render.fps`` / ``render.seed`` (``FluidState::new``, fluid.rs:109), reports
``particle_count`` and ``reservation`` (fluid.rs:104-106, 230-232) and is advanced to a frame
with ``advance`` (``fluid_advance``'s step loop, fluid.rs:265-306).  Everything runs in the
HIP library on the GPU; there is no fallback: if the library is missing the import fails.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import time
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpxlflip.so")

OK = 0
ERR_INVALID, ERR_OOM, ERR_DEVICE, ERR_INTERNAL = -1, -2, -3, -4

SOLVERS = {"rbgs": 0, "cg": 1, "mixed": 2}

FIELDS = {
    "vel": (0, np.float32, "faces"),
    "vel_old": (1, np.float32, "faces"),
    "acc": (2, np.int32, "faces"),
    "wacc": (3, np.int32, "faces"),
    "count": (4, np.uint32, "cells"),
    "type": (5, np.uint32, "cells"),
    "pressure": (6, np.float32, "cells"),
    "divergence": (7, np.float32, "cells"),
    "pressure64": (8, np.float64, "cells"),
    # surface nets (after FluidState.surface)
    "density": (9, np.int32, "snodes"),
    "phi": (10, np.float32, "snodes"),
    "cell_vert": (11, np.uint32, "scells"),
}
WRITABLE = ("vel", "vel_old", "pressure")


class FluidError(RuntimeError):
    """A status code other than PXLFLIP_OK, with the library's message."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class Opts(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32), ("solver", C.c_int32), ("cg_iterations", C.c_uint32),
        ("inner_iterations", C.c_uint32), ("max_outer", C.c_uint32), ("tolerance", C.c_double),
        ("device", C.c_int32), ("scene_index", C.c_uint32), ("use_graph", C.c_int32),
        ("solver_workgroups", C.c_uint32), ("max_cells", C.c_uint32), ("inner_max", C.c_uint32),
    ]


class Info(C.Structure):
    _fields_ = [
        ("n", C.c_uint32 * 3), ("h", C.c_float), ("dt", C.c_float), ("flip", C.c_float),
        ("inv_h", C.c_float), ("origin", C.c_float * 3), ("gravity", C.c_float * 3),
        ("n_particles", C.c_uint32), ("substeps", C.c_uint32), ("pressure_iterations", C.c_uint32),
        ("n_faces", C.c_uint64), ("n_cells", C.c_uint64), ("device_bytes", C.c_uint64),
    ]

    def as_dict(self):
        return {
            "n": tuple(int(v) for v in self.n), "h": np.float32(self.h), "dt": np.float32(self.dt),
            "flip": np.float32(self.flip), "inv_h": np.float32(self.inv_h),
            "origin": tuple(np.float32(v) for v in self.origin),
            "gravity": tuple(np.float32(v) for v in self.gravity),
            "n_particles": int(self.n_particles), "substeps": int(self.substeps),
            "pressure_iterations": int(self.pressure_iterations), "n_faces": int(self.n_faces),
            "n_cells": int(self.n_cells), "device_bytes": int(self.device_bytes),
        }


class SceneInfo(C.Structure):
    _fields_ = [
        ("version", C.c_uint32), ("n_nodes", C.c_uint32), ("n_simulations", C.c_uint32), ("n_fluids", C.c_uint32),
        ("fps", C.c_double), ("seed", C.c_uint32), ("frame_start", C.c_int32), ("frame_end", C.c_int32),
        ("pad", C.c_uint32),
    ]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_uint32), ("total_ms", C.c_double),
                ("alg_bytes", C.c_double)]


class SolveStats(C.Structure):
    _fields_ = [("rel_residual", C.c_double), ("inner_iterations", C.c_uint32),
                ("outer_iterations", C.c_uint32)]


class ServerOpts(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32), ("device", C.c_int32), ("budget_bytes", C.c_uint64),
        ("max_sims", C.c_uint32), ("inner_iterations", C.c_uint32), ("max_outer", C.c_uint32),
        ("tolerance", C.c_double), ("solver_workgroups", C.c_uint32), ("use_graph", C.c_int32),
        ("profile_every", C.c_uint32),
    ]


class RequestInfo(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("steps", C.c_uint32), ("steps_done", C.c_uint32), ("n_particles", C.c_uint32),
        ("device_bytes", C.c_uint64), ("queue_ms", C.c_double), ("latency_ms", C.c_double),
        ("rel_residual", C.c_double), ("max_outer", C.c_uint32), ("pad", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}
        d["status"] = REQUEST_STATUS[self.status]
        return d


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
        ("cancelled", C.c_uint64), ("queued", C.c_uint64), ("running", C.c_uint64),
        ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double),
        ("particle_steps", C.c_double), ("span_s", C.c_double), ("throughput", C.c_double),
        ("busy_s", C.c_double), ("batch_steps", C.c_uint64), ("mean_batch", C.c_double),
        ("bytes_in_use", C.c_uint64), ("peak_bytes", C.c_uint64), ("admit_s", C.c_double),
        ("rebuilds", C.c_uint64), ("graph_captures", C.c_uint64), ("capture_s", C.c_double),
        ("retire_s", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class DeviceBuffer:
    """Device memory from the HIP runtime the library uses (for the device entry points, e.g.
    the renderer-side vertex / index buffers of FluidState.surface_device)."""

    _hip = None

    def __init__(self, nbytes: int):
        if DeviceBuffer._hip is None:
            path = "/opt/rocm/lib/libamdhip64.so"
            h = C.CDLL(path if os.path.exists(path) else "libamdhip64.so")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipFree.argtypes = [C.c_void_p]
            DeviceBuffer._hip = h
        self.nbytes = int(nbytes)
        self._p = C.c_void_p()
        if DeviceBuffer._hip.hipMalloc(C.byref(self._p), max(self.nbytes, 4)) != 0:
            raise MemoryError(f"hipMalloc of {nbytes} bytes failed")

    @property
    def ptr(self) -> int:
        return int(self._p.value)

    def upload(self, host: np.ndarray):
        assert host.nbytes <= self.nbytes
        if DeviceBuffer._hip.hipMemcpy(self._p, host.ctypes.data, host.nbytes, 1) != 0:
            raise FluidError(ERR_DEVICE, "hipMemcpy to the device failed")

    def copy_from(self, other: "DeviceBuffer", nbytes: int = None):
        """Device-to-device copy of nbytes (default all of other) into this buffer."""
        n = other.nbytes if nbytes is None else int(nbytes)
        assert n <= self.nbytes and n <= other.nbytes
        if DeviceBuffer._hip.hipMemcpy(self._p, other._p, n, 3) != 0:
            raise FluidError(ERR_DEVICE, "hipMemcpy device to device failed")

    def download(self, out: np.ndarray) -> np.ndarray:
        assert out.nbytes <= self.nbytes
        if DeviceBuffer._hip.hipMemcpy(out.ctypes.data, self._p, out.nbytes, 2) != 0:
            raise FluidError(ERR_DEVICE, "hipMemcpy from the device failed")
        return out

    def free(self):
        if self._p:
            DeviceBuffer._hip.hipFree(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class ImportedMemory:
    """An allocation of another API (e.g. the Vulkan memory block the renderer suballocates its
    vertex and index buffers from) imported once from an exported file descriptor; map() gives
    device memory for byte ranges of it."""

    def __init__(self, fd: int, nbytes: int, dedicated: bool = False, device: int = 0):
        self._lib = load_library()
        self._h = C.c_void_p()
        rc = self._lib.pxlflip_import_memory_fd(int(fd), int(nbytes), 1 if dedicated else 0, int(device),
                                                C.byref(self._h))
        if rc != 0:
            raise FluidError(rc, self._lib.pxlflip_last_error(None).decode())
        self.nbytes = int(nbytes)

    def map(self, offset: int, nbytes: int) -> int:
        p = C.c_void_p()
        rc = self._lib.pxlflip_external_map(self._h, int(offset), int(nbytes), C.byref(p))
        if rc != 0:
            raise FluidError(rc, self._lib.pxlflip_last_error(None).decode())
        return int(p.value)

    def release(self):
        if self._h:
            self._lib.pxlflip_release_buffer(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class ImportedBuffer:
    """A buffer of another API (e.g. a Vulkan vertex buffer) imported from an exported file
    descriptor; .ptr is device memory for FluidState.surface_device."""

    def __init__(self, fd: int, nbytes: int, device: int = 0):
        self._lib = load_library()
        self._h = C.c_void_p()
        p = C.c_void_p()
        rc = self._lib.pxlflip_import_buffer_fd(int(fd), int(nbytes), int(device), C.byref(p), C.byref(self._h))
        if rc != 0:
            raise FluidError(rc, self._lib.pxlflip_last_error(None).decode())
        self.ptr = int(p.value)
        self.nbytes = int(nbytes)

    def release(self):
        if self._h:
            self._lib.pxlflip_release_buffer(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class AgentGrid:
    """Neighbour grid of an agent simulation (shaders/sim.slang:27-104): agents grouped by the
    hash of their grid cell."""

    def __init__(self, max_agents: int, device: int = 0):
        self._lib = load_library()
        self._h = C.c_void_p()
        rc = self._lib.pxlgrid_create(int(max_agents), int(device), C.byref(self._h))
        if rc != 0:
            raise FluidError(rc, self._lib.pxlgrid_last_error(None).decode())
        self.max_agents = int(max_agents)

    @staticmethod
    def table_size(count: int) -> int:
        return int(load_library().pxlgrid_table_size(int(count)))

    def _check(self, rc):
        if rc != 0:
            raise FluidError(rc, self._lib.pxlgrid_last_error(self._h).decode())

    def build(self, agents: np.ndarray, cell_size: float, table_size: int = None):
        """agents[N, 16] float32 (64-byte records) -> (offsets, counts, sorted)."""
        a = np.ascontiguousarray(agents, np.float32)
        assert a.ndim == 2 and a.shape[1] == 16
        n = len(a)
        t = self.table_size(n) if table_size is None else int(table_size)
        off = np.zeros(t, np.uint32)
        cnt = np.zeros(t, np.uint32)
        srt = np.zeros(max(n, 1), np.uint32)
        self._check(self._lib.pxlgrid_build_host(self._h, a.ctypes.data, n, float(np.float32(cell_size)), t,
                                                 off.ctypes.data, cnt.ctypes.data, srt.ctypes.data))
        return off, cnt, srt[:n]

    def build_device(self, agents_ptr: int, count: int, cell_size: float, table_size: int, offsets_ptr: int,
                     counts_ptr: int, sorted_ptr: int) -> float:
        """The same on device buffers; returns the kernels' time in ms."""
        self._check(self._lib.pxlgrid_build(self._h, C.c_void_p(agents_ptr), int(count), float(np.float32(cell_size)),
                                            int(table_size), C.c_void_p(offsets_ptr), C.c_void_p(counts_ptr),
                                            C.c_void_p(sorted_ptr)))
        return float(self._lib.pxlgrid_last_ms(self._h))

    def close(self):
        if self._h:
            self._lib.pxlgrid_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SurfaceInfo(C.Structure):
    _fields_ = [
        ("sx", C.c_uint32), ("sy", C.c_uint32), ("sz", C.c_uint32), ("max_triangles", C.c_uint32),
        ("vertices", C.c_uint32), ("triangles", C.c_uint32), ("vertices_total", C.c_uint32),
        ("triangles_total", C.c_uint32), ("overflow", C.c_uint32),
        ("sh", C.c_float), ("radius", C.c_float), ("iso", C.c_float),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


REQUEST_STATUS = {0: "queued", 1: "running", 2: "done", 3: "failed", 4: "cancelled"}


# Exported symbols with their ctypes signatures (must match include/pxlflip.h).
_H = C.c_void_p
_SIGNATURES = {
    "pxlflip_default_opts": (C.c_int, [C.POINTER(Opts)]),
    "pxlflip_check_json": (C.c_int, [C.c_char_p, C.c_double, C.c_uint32, C.POINTER(Opts), C.POINTER(Info)]),
    "pxlflip_scene_check": (C.c_int, [C.c_char_p, C.POINTER(Opts), C.POINTER(SceneInfo), C.c_char_p, C.c_size_t]),
    "pxlflip_create_from_scene": (C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(Opts), C.POINTER(_H)]),
    "pxlflip_seed_particles": (C.c_int, [C.c_char_p, C.c_uint32, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "pxlflip_create_from_json": (C.c_int, [C.c_char_p, C.c_double, C.c_uint32, C.POINTER(Opts), C.POINTER(_H)]),
    "pxlflip_destroy": (None, [_H]),
    "pxlflip_surface": (C.c_int, [_H, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.POINTER(SurfaceInfo)]),
    "pxlflip_surface_host": (C.c_int, [_H, C.c_void_p, C.c_void_p, C.POINTER(SurfaceInfo)]),
    "pxlflip_last_substeps": (C.c_int, [_H, C.POINTER(C.c_uint32)]),
    "pxlflip_import_buffer_fd": (C.c_int, [C.c_int, C.c_uint64, C.c_int, C.POINTER(C.c_void_p), C.POINTER(_H)]),
    "pxlflip_import_memory_fd": (C.c_int, [C.c_int, C.c_uint64, C.c_uint32, C.c_int, C.POINTER(_H)]),
    "pxlflip_external_map": (C.c_int, [_H, C.c_uint64, C.c_uint64, C.POINTER(C.c_void_p)]),
    "pxlflip_release_buffer": (None, [_H]),
    "pxlgrid_table_size": (C.c_uint32, [C.c_uint32]),
    "pxlgrid_create": (C.c_int, [C.c_uint32, C.c_int, C.POINTER(_H)]),
    "pxlgrid_build": (C.c_int, [_H, C.c_void_p, C.c_uint32, C.c_float, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pxlgrid_build_host": (C.c_int, [_H, C.c_void_p, C.c_uint32, C.c_float, C.c_uint32, C.c_void_p, C.c_void_p,
                                     C.c_void_p]),
    "pxlgrid_build_static": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_uint32,
                                       C.c_void_p, C.c_void_p, C.c_void_p]),
    "pxlgrid_last_ms": (C.c_double, [_H]),
    "pxlgrid_last_error": (C.c_char_p, [_H]),
    "pxlgrid_destroy": (None, [_H]),
    "pxlflip_advance_to_frame": (C.c_int, [_H, C.c_double]),
    "pxlflip_step": (C.c_int, [_H, C.c_uint32]),
    "pxlflip_reset": (C.c_int, [_H]),
    "pxlflip_particle_count": (C.c_uint32, [_H]),
    "pxlflip_steps_done": (C.c_uint64, [_H]),
    "pxlflip_reservation": (C.c_int, [_H, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "pxlflip_get_info": (C.c_int, [_H, C.POINTER(Info)]),
    "pxlflip_read_particles": (C.c_int, [_H, C.c_void_p]),
    "pxlflip_write_particles": (C.c_int, [_H, C.c_void_p]),
    "pxlflip_read_field": (C.c_int, [_H, C.c_int, C.c_void_p]),
    "pxlflip_write_field": (C.c_int, [_H, C.c_int, C.c_void_p]),
    "pxlflip_read_bins": (C.c_int, [_H, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pxlflip_grid_phase": (C.c_int, [_H]),
    "pxlflip_last_solve": (C.c_int, [_H, C.POINTER(SolveStats)]),
    "pxlflip_profile_steps": (C.c_int, [_H, C.c_uint32, C.POINTER(KernelTime), C.c_uint32, C.POINTER(C.c_uint32)]),
    "pxlflip_profile_surface": (C.c_int, [_H, C.c_uint32, C.POINTER(KernelTime), C.c_uint32, C.POINTER(C.c_uint32)]),
    "pxlflip_autotune": (C.c_int, [_H, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_double)]),
    "pxlflip_batch_get_grid": (C.c_int, [C.c_int, C.POINTER(C.c_uint32)]),
    "pxlflip_batch_get_shape": (C.c_int, [C.c_int, C.POINTER(C.c_uint32)]),
    "pxlflip_batch_set_shape": (C.c_int, [C.c_int, C.c_uint32]),
    "pxlflip_batch_set_grid": (C.c_int, [C.c_int, C.c_uint32]),
    "pxlflip_last_error": (C.c_char_p, [_H]),
    "pxlflip_server_default_opts": (C.c_int, [C.POINTER(ServerOpts)]),
    "pxlflip_server_create": (C.c_int, [C.POINTER(ServerOpts), C.POINTER(_H)]),
    "pxlflip_server_submit": (C.c_int, [_H, C.c_char_p, C.c_double, C.c_uint32, C.c_uint32, C.c_int32,
                                        C.POINTER(C.c_uint64)]),
    "pxlflip_server_poll": (C.c_int, [_H, C.c_uint64, C.POINTER(RequestInfo)]),
    "pxlflip_server_wait": (C.c_int, [_H, C.c_uint64, C.c_double, C.POINTER(RequestInfo)]),
    "pxlflip_server_drain": (C.c_int, [_H, C.c_double]),
    "pxlflip_server_cancel": (C.c_int, [_H, C.c_uint64]),
    "pxlflip_server_read_particles": (C.c_int, [_H, C.c_uint64, C.c_void_p, C.c_uint64]),
    "pxlflip_server_release": (C.c_int, [_H, C.c_uint64]),
    "pxlflip_server_get_stats": (C.c_int, [_H, C.POINTER(ServerStats)]),
    "pxlflip_server_kernel_stats": (C.c_int, [_H, C.POINTER(KernelTime), C.c_uint32, C.POINTER(C.c_uint32)]),
    "pxlflip_server_last_error": (C.c_char_p, [_H]),
    "pxlflip_server_destroy": (None, [_H]),
    "pxlflip_batch_create": (C.c_int, [C.POINTER(C.c_char_p), C.POINTER(C.c_uint32), C.c_uint32, C.c_double,
                                       C.POINTER(ServerOpts), C.POINTER(_H)]),
    "pxlflip_batch_step": (C.c_int, [_H, C.c_uint32]),
    "pxlflip_batch_profile_steps": (C.c_int, [_H, C.c_uint32, C.POINTER(KernelTime), C.c_uint32,
                                              C.POINTER(C.c_uint32)]),
    "pxlflip_batch_get_info": (C.c_int, [_H, C.c_uint32, C.POINTER(Info)]),
    "pxlflip_batch_read_particles": (C.c_int, [_H, C.c_uint32, C.c_void_p]),
    "pxlflip_batch_last_solve": (C.c_int, [_H, C.c_uint32, C.POINTER(SolveStats)]),
    "pxlflip_batch_read_field": (C.c_int, [_H, C.c_uint32, C.c_int, C.c_void_p]),
    "pxlflip_batch_read_bins": (C.c_int, [_H, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pxlflip_batch_last_error": (C.c_char_p, [_H]),
    "pxlflip_batch_destroy": (None, [_H]),
}

_lib = None


def load_library(path: Optional[str] = None):
    """Loads the HIP library; raises ImportError when it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(f"{p} is missing: build it with `python -m pxlengine_amd.build` "
                          "(the FLIP path has no CPU fallback)")
    lib = C.CDLL(p)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _err(handle=None) -> str:
    msg = load_library().pxlflip_last_error(handle)
    return msg.decode() if msg else ""


def _check(rc: int, handle=None):
    if rc != OK:
        raise FluidError(rc, _err(handle))


def _json(fluid) -> bytes:
    if isinstance(fluid, (bytes, bytearray)):
        return bytes(fluid)
    if isinstance(fluid, str):
        return fluid.encode()
    return json.dumps(fluid).encode()


def make_opts(solver="rbgs", cg_iterations=0, inner_iterations=9, max_outer=50, tolerance=1e-6,
              device=0, scene_index=0, use_graph=True, solver_workgroups=0, max_cells=0, inner_max=0) -> Opts:
    o = Opts()
    load_library().pxlflip_default_opts(C.byref(o))
    o.solver = SOLVERS[solver] if isinstance(solver, str) else int(solver)
    o.cg_iterations = cg_iterations
    o.inner_iterations = inner_iterations
    o.max_outer = max_outer
    o.tolerance = tolerance
    o.device = device
    o.scene_index = scene_index
    o.use_graph = 1 if use_graph else 0
    o.solver_workgroups = solver_workgroups
    o.max_cells = max_cells
    o.inner_max = inner_max
    return o


def check(fluid, fps: float = 24.0, seed: int = 0, **opts) -> dict:
    """Parse + validate a fluid block without touching the GPU (Scene::validate's fluid
    checks, crates/pxl-scene/src/lib.rs:710-735, plus Fluid::grid and the particle count)."""
    info = Info()
    o = make_opts(**opts)
    _check(load_library().pxlflip_check_json(_json(fluid), fps, seed, C.byref(o), C.byref(info)))
    return info.as_dict()


def check_scene(scene, **opts) -> dict:
    """Parse + validate a whole scene document for the fluid path without touching the GPU
    (Scene::from_json / validate, crates/pxl-scene/src/lib.rs:646-745) and the `pxl check`
    summary line (crates/pxl-cli/src/check.rs:98-104)."""
    info = SceneInfo()
    line = C.create_string_buffer(256)
    o = make_opts(**opts)
    _check(load_library().pxlflip_scene_check(_json(scene), C.byref(o), C.byref(info), line, 256))
    return {"version": int(info.version), "nodes": int(info.n_nodes), "simulations": int(info.n_simulations),
            "fluids": int(info.n_fluids), "fps": float(info.fps), "seed": int(info.seed),
            "frames": (int(info.frame_start), int(info.frame_end)), "line": line.value.decode()}


def seed_particles(fluid, seed: int = 0) -> np.ndarray:
    """Initial particles as (P, 8) float32 records (fluid.rs:123-148), computed on the host."""
    L = load_library()
    cnt = C.c_uint64()
    _check(L.pxlflip_seed_particles(_json(fluid), seed, None, 0, C.byref(cnt)))
    out = np.zeros((cnt.value, 8), np.float32)
    _check(L.pxlflip_seed_particles(_json(fluid), seed, out.ctypes.data, cnt.value, C.byref(cnt)))
    return out


class FluidState:
    """One simulated liquid resident in HBM (crates/pxl-render/src/fluid.rs:74-94)."""

    def __init__(self, fluid, fps: float = 24.0, seed: int = 0, **opts):
        self._lib = load_library()
        self._h = C.c_void_p()
        o = make_opts(**opts)
        _check(self._lib.pxlflip_create_from_json(_json(fluid), fps, seed, C.byref(o), C.byref(self._h)))
        self._finish(fps, seed)

    @classmethod
    def from_scene(cls, scene, fluid_index: int = 0, **opts) -> "FluidState":
        """fluids[fluid_index] of a whole scene document, with its render.fps and render.seed."""
        self = cls.__new__(cls)
        self._lib = load_library()
        self._h = C.c_void_p()
        o = make_opts(**opts)
        _check(self._lib.pxlflip_create_from_scene(_json(scene), int(fluid_index), C.byref(o), C.byref(self._h)))
        sc = check_scene(scene, **opts)
        self._finish(sc["fps"], sc["seed"])
        return self

    def _finish(self, fps, seed):
        info = Info()
        _check(self._lib.pxlflip_get_info(self._h, C.byref(info)), self._h)
        self.info = info.as_dict()
        self.fps = fps
        self.seed = seed

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.pxlflip_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _call(self, fn, *args):
        if not self._h.value:
            raise FluidError(ERR_INVALID, "handle is closed")
        _check(fn(self._h, *args), self._h)

    # -- reference interface
    def particle_count(self) -> int:
        return int(self._lib.pxlflip_particle_count(self._h))

    def reservation(self):
        v, i = C.c_uint32(), C.c_uint32()
        self._call(self._lib.pxlflip_reservation, C.byref(v), C.byref(i))
        return int(v.value), int(i.value)

    def advance(self, frame: float):
        self._call(self._lib.pxlflip_advance_to_frame, float(frame))

    @property
    def steps_done(self) -> int:
        return int(self._lib.pxlflip_steps_done(self._h))

    # -- build extensions
    def step(self, n: int = 1):
        self._call(self._lib.pxlflip_step, int(n))

    def reset(self):
        self._call(self._lib.pxlflip_reset)

    def grid_phase(self):
        self._call(self._lib.pxlflip_grid_phase)

    def read_particles(self) -> np.ndarray:
        out = np.empty((self.info["n_particles"], 8), np.float32)
        self._call(self._lib.pxlflip_read_particles, out.ctypes.data)
        return out

    def write_particles(self, aos8: np.ndarray):
        a = np.ascontiguousarray(aos8, np.float32)
        if a.shape != (self.info["n_particles"], 8):
            raise ValueError(f"expected ({self.info['n_particles']}, 8) records, got {a.shape}")
        self._call(self._lib.pxlflip_write_particles, a.ctypes.data)

    def last_substeps(self) -> int:
        """Substeps of the last step under the CFL cap (1 without it)."""
        k = C.c_uint32()
        self._call(self._lib.pxlflip_last_substeps, C.byref(k))
        return int(k.value)

    def surface(self):
        """Surface nets of the current particles (fluid.rs:308-330) -> (verts[V, 8] as {pos, 0,
        normal, 0}, indices[T, 3], info)."""
        cap = self.reservation()[0]
        verts = np.zeros((max(cap, 1), 8), np.float32)
        idx = np.zeros((max(cap, 1), 3), np.uint32)
        info = SurfaceInfo()
        self._call(self._lib.pxlflip_surface_host, verts.ctypes.data, idx.ctypes.data, C.byref(info))
        self._surf = info.as_dict()
        return verts[:info.vertices], idx[:info.triangles], self._surf

    def surface_device(self, verts_ptr: int, vbase: int, indices_ptr: int, ibase: int) -> dict:
        """The same into device buffers (e.g. torch tensors' data_ptr()), as the renderer binds them."""
        info = SurfaceInfo()
        self._call(self._lib.pxlflip_surface, C.c_void_p(verts_ptr), vbase, C.c_void_p(indices_ptr), ibase,
                   C.byref(info))
        self._surf = info.as_dict()
        return self._surf

    def read_field(self, name: str) -> np.ndarray:
        code, dtype, kind = FIELDS[name]
        if kind in ("snodes", "scells"):
            sf = getattr(self, "_surf", None)
            if sf is None:
                raise FluidError("no surface pass has run on this handle")
            n = ((sf["sx"] + 1) * (sf["sy"] + 1) * (sf["sz"] + 1) if kind == "snodes"
                 else sf["sx"] * sf["sy"] * sf["sz"])
        else:
            n = self.info["n_faces"] if kind == "faces" else self.info["n_cells"]
        out = np.empty(n, dtype)
        self._call(self._lib.pxlflip_read_field, code, out.ctypes.data)
        return out

    def write_field(self, name: str, values: np.ndarray):
        if name not in WRITABLE:
            raise ValueError(f"{name} is not writable")
        code, dtype, kind = FIELDS[name]
        n = self.info["n_faces"] if kind == "faces" else self.info["n_cells"]
        a = np.ascontiguousarray(values, dtype)
        if a.shape != (n,):
            raise ValueError(f"expected {n} values, got {a.shape}")
        self._call(self._lib.pxlflip_write_field, code, a.ctypes.data)

    def read_bins(self):
        P, Cc = self.info["n_particles"], self.info["n_cells"]
        cell_of = np.empty(P, np.uint32)
        counts = np.empty(Cc, np.uint32)
        offsets = np.empty(Cc, np.uint32)
        ids = np.empty(P, np.uint32)
        self._call(self._lib.pxlflip_read_bins, cell_of.ctypes.data, counts.ctypes.data, offsets.ctypes.data,
                   ids.ctypes.data)
        return cell_of, counts, offsets, ids

    def profile_steps(self, n: int = 1) -> list:
        """Runs n steps with every launch timed by HIP events; [(name, launches, total_ms)]."""
        arr = (KernelTime * 64)()
        cnt = C.c_uint32()
        self._call(self._lib.pxlflip_profile_steps, int(n), arr, 64, C.byref(cnt))
        return [(arr[k].name.decode(), int(arr[k].launches), float(arr[k].total_ms))
                for k in range(min(cnt.value, 64))]

    def profile_surface(self, n: int = 1) -> list:
        """Runs n surface passes with every stage timed by HIP events; [(name, launches, total_ms)]."""
        arr = (KernelTime * 16)()
        cnt = C.c_uint32()
        self._call(self._lib.pxlflip_profile_surface, int(n), arr, 16, C.byref(cnt))
        return [(arr[k].name.decode(), int(arr[k].launches), float(arr[k].total_ms))
                for k in range(min(cnt.value, 16))]

    def autotune(self, steps: int = 2) -> dict:
        """Picks the solver workgroup count by timing `steps` steps per candidate; state unchanged."""
        g, ms = C.c_uint32(), C.c_double()
        self._call(self._lib.pxlflip_autotune, int(steps), C.byref(g), C.byref(ms))
        return {"solver_workgroups": int(g.value), "ms_per_step": float(ms.value)}

    def last_solve(self) -> dict:
        st = SolveStats()
        self._call(self._lib.pxlflip_last_solve, C.byref(st))
        return {"rel_residual": st.rel_residual, "inner_iterations": int(st.inner_iterations),
                "outer_iterations": int(st.outer_iterations)}


def make_server_opts(device=0, budget_bytes=200_000_000_000, max_sims=256, inner_iterations=9, max_outer=50,
                     tolerance=1e-6, solver_workgroups=0, use_graph=True, profile_every=0) -> ServerOpts:
    o = ServerOpts()
    load_library().pxlflip_server_default_opts(C.byref(o))
    o.device = device
    o.budget_bytes = int(budget_bytes)
    o.max_sims = max_sims
    o.inner_iterations = inner_iterations
    o.max_outer = max_outer
    o.tolerance = tolerance
    o.solver_workgroups = solver_workgroups
    o.use_graph = 1 if use_graph else 0
    o.profile_every = int(profile_every)
    return o


class Server:
    """Continuous-batching server of independent liquids (pxlflip_server_* in pxlflip.h).

    ``submit`` validates a fluid block and queues ``steps`` steps of it; a device thread admits
    queued requests while their HBM fits ``budget_bytes``, steps every resident simulation
    together with batched launches (mixed-precision solve), and retires each after its last
    step.  Each simulation's result equals its result alone.
    """

    def __init__(self, **opts):
        self._lib = load_library()
        self._h = C.c_void_p()
        o = make_server_opts(**opts)
        rc = self._lib.pxlflip_server_create(C.byref(o), C.byref(self._h))
        if rc != OK:
            raise FluidError(rc, self._lib.pxlflip_server_last_error(None).decode())

    def _check(self, rc: int):
        if rc != OK:
            msg = self._lib.pxlflip_server_last_error(self._h)
            raise FluidError(rc, msg.decode() if msg else "")

    def submit(self, fluid, fps: float = 24.0, seed: int = 0, steps: int = 1, keep_particles: bool = False) -> int:
        rid = C.c_uint64()
        self._check(self._lib.pxlflip_server_submit(self._h, _json(fluid), fps, seed, steps,
                                                    1 if keep_particles else 0, C.byref(rid)))
        return int(rid.value)

    def poll(self, rid: int) -> dict:
        info = RequestInfo()
        self._check(self._lib.pxlflip_server_poll(self._h, rid, C.byref(info)))
        return info.as_dict()

    def wait(self, rid: int, timeout: float = -1.0) -> dict:
        info = RequestInfo()
        self._check(self._lib.pxlflip_server_wait(self._h, rid, timeout, C.byref(info)))
        return info.as_dict()

    def drain(self, timeout: float = -1.0):
        self._check(self._lib.pxlflip_server_drain(self._h, timeout))

    def cancel(self, rid: int):
        self._check(self._lib.pxlflip_server_cancel(self._h, rid))

    def read_particles(self, rid: int) -> np.ndarray:
        n = self.poll(rid)["n_particles"]
        out = np.zeros((n, 8), np.float32)
        self._check(self._lib.pxlflip_server_read_particles(self._h, rid, out.ctypes.data, n))
        return out

    def release(self, rid: int):
        self._check(self._lib.pxlflip_server_release(self._h, rid))

    def stats(self) -> dict:
        st = ServerStats()
        self._check(self._lib.pxlflip_server_get_stats(self._h, C.byref(st)))
        return st.as_dict()

    def kernel_stats(self) -> list:
        """Per-kernel totals of the profiled batch steps (profile_every):
        [{"kernel", "launches", "total_ms", "alg_bytes"}]."""
        arr = (KernelTime * 128)()
        cnt = C.c_uint32()
        self._check(self._lib.pxlflip_server_kernel_stats(self._h, arr, 128, C.byref(cnt)))
        return [{"kernel": arr[k].name.decode(), "launches": int(arr[k].launches),
                 "total_ms": float(arr[k].total_ms), "alg_bytes": float(arr[k].alg_bytes)}
                for k in range(min(cnt.value, 128))]

    def close(self):
        if self._h:
            self._lib.pxlflip_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """A fixed set of liquids advanced in lock step by the batched launches of the server
    (pxlflip_batch_* in pxlflip.h), mixed-precision solve; no serving loop."""

    def __init__(self, fluids, fps: float = 24.0, seeds=None, **opts):
        self._lib = load_library()
        self._h = C.c_void_p()
        n = len(fluids)
        arr = (C.c_char_p * n)(*[_json(f) for f in fluids])
        sd = (C.c_uint32 * n)(*(seeds if seeds is not None else [0] * n))
        o = make_server_opts(**opts)
        rc = self._lib.pxlflip_batch_create(arr, sd, n, fps, C.byref(o), C.byref(self._h))
        if rc != OK:
            raise FluidError(rc, self._lib.pxlflip_batch_last_error(None).decode())
        self.n = n
        self.infos = []
        for i in range(n):
            info = Info()
            self._check(self._lib.pxlflip_batch_get_info(self._h, i, C.byref(info)))
            self.infos.append(info.as_dict())

    def _check(self, rc: int):
        if rc != OK:
            msg = self._lib.pxlflip_batch_last_error(self._h)
            raise FluidError(rc, msg.decode() if msg else "")

    def step(self, n: int = 1):
        self._check(self._lib.pxlflip_batch_step(self._h, int(n)))

    def profile_steps(self, n: int = 1) -> list:
        """Runs n steps with every launch timed by HIP events:
        [{"kernel", "launches", "total_ms", "alg_bytes"}] in first-launch order."""
        arr = (KernelTime * 128)()
        cnt = C.c_uint32()
        self._check(self._lib.pxlflip_batch_profile_steps(self._h, int(n), arr, 128, C.byref(cnt)))
        return [{"kernel": arr[k].name.decode(), "launches": int(arr[k].launches),
                 "total_ms": float(arr[k].total_ms), "alg_bytes": float(arr[k].alg_bytes)}
                for k in range(min(cnt.value, 128))]

    def read_particles(self, index: int) -> np.ndarray:
        out = np.zeros((self.infos[index]["n_particles"], 8), np.float32)
        self._check(self._lib.pxlflip_batch_read_particles(self._h, index, out.ctypes.data))
        return out

    def read_field(self, index: int, name: str) -> np.ndarray:
        """Member `index`'s grid field (the FIELDS names; surface fields are not kept per member)."""
        code, dtype, kind = FIELDS[name]
        if kind not in ("faces", "cells"):
            raise ValueError(f"{name} is not a grid field")
        out = np.empty(self.infos[index]["n_faces" if kind == "faces" else "n_cells"], dtype)
        self._check(self._lib.pxlflip_batch_read_field(self._h, index, code, out.ctypes.data))
        return out

    def read_bins(self, index: int):
        """Counting sort of member `index`'s current positions: (cell_of, counts, offsets, ids)."""
        P, Cc = self.infos[index]["n_particles"], self.infos[index]["n_cells"]
        out = (np.empty(P, np.uint32), np.empty(Cc, np.uint32), np.empty(Cc, np.uint32), np.empty(P, np.uint32))
        self._check(self._lib.pxlflip_batch_read_bins(self._h, index, *(a.ctypes.data for a in out)))
        return out

    def last_solve(self, index: int) -> dict:
        st = SolveStats()
        self._check(self._lib.pxlflip_batch_last_solve(self._h, index, C.byref(st)))
        return {"rel_residual": st.rel_residual, "inner_iterations": int(st.inner_iterations),
                "outer_iterations": int(st.outer_iterations)}

    def close(self):
        if self._h:
            self._lib.pxlflip_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- batched launch grids and their autotuner (pxlflip_batch_get_grid / _set_grid)

# PXLFLIP_GRID_* of pxlflip.h, in order
GRID_FAMILIES = ("band", "particle", "tile", "mg_setup", "mg_list", "mg_tail", "solver")
# PXLFLIP_SHAPE_* of pxlflip.h, in order, with the workgroup sizes compiled for each (default first)
SHAPE_FAMILIES = {"tile_lanes": (512, 1024), "band_lanes": (256, 512)}


def batch_shapes() -> dict:
    """The workgroup size of every batched launch family compiled in more than one size."""
    lib = load_library()
    out = {}
    for i, name in enumerate(SHAPE_FAMILIES):
        v = C.c_uint32()
        _check(lib.pxlflip_batch_get_shape(i, C.byref(v)))
        out[name] = int(v.value)
    return out


def set_batch_shapes(shapes: dict):
    """Sets the workgroup sizes of the named shape families (one of their compiled sizes)."""
    lib = load_library()
    names = list(SHAPE_FAMILIES)
    for name, v in shapes.items():
        if name not in SHAPE_FAMILIES:
            raise KeyError(f"unknown batch shape family {name!r} (one of {', '.join(names)})")
        rc = lib.pxlflip_batch_set_shape(names.index(name), int(v))
        if rc != OK:
            raise FluidError(rc, f"batch shape {name}: {v} is not one of {SHAPE_FAMILIES[name]}")


def batch_grids() -> dict:
    """The persistent-grid maximum of every batched launch family (process-wide)."""
    lib = load_library()
    out = {}
    for i, name in enumerate(GRID_FAMILIES):
        v = C.c_uint32()
        _check(lib.pxlflip_batch_get_grid(i, C.byref(v)))
        out[name] = int(v.value)
    return out


def set_batch_grids(grids: dict):
    """Sets the maxima of the named families (multiples of 8 in [8, 16384]); they apply to
    batch graphs captured afterwards.  Unknown names raise KeyError, bad values FluidError."""
    lib = load_library()
    for name, v in grids.items():
        i = GRID_FAMILIES.index(name) if name in GRID_FAMILIES else None
        if i is None:
            raise KeyError(f"unknown batch grid family {name!r} (one of {', '.join(GRID_FAMILIES)})")
        rc = lib.pxlflip_batch_set_grid(i, int(v))
        if rc != OK:
            raise FluidError(rc, f"batch grid {name}: {v} is not a multiple of 8 in [8, 16384]")


def save_batch_grids(path: str, grids: Optional[dict] = None, shapes: Optional[dict] = None, **extra):
    """Writes the grids and workgroup sizes (default: the current ones) to a JSON cache file."""
    doc = {"grids": dict(grids if grids is not None else batch_grids()),
           "shapes": dict(shapes if shapes is not None else batch_shapes())}
    doc.update(extra)
    with open(path, "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)


def load_batch_grids(path: str) -> dict:
    """Applies the grids (and workgroup sizes, when present) of a JSON cache file written by
    save_batch_grids / autotune_batch; returns the grids."""
    with open(path) as f:
        doc = json.load(f)
    set_batch_grids(doc["grids"])
    if "shapes" in doc:
        set_batch_shapes(doc["shapes"])
    return doc["grids"]


def _time_batch(fluids, fps, steps, warmup, opts) -> float:
    with Batch(fluids, fps, **opts) as b:
        b.step(warmup)
        t0 = time.perf_counter()
        b.step(steps)
        return (time.perf_counter() - t0) / steps


def autotune_batch(fluids, fps: float = 24.0, steps: int = 3, warmup: int = 1, families=None,
                   factors=(0.5, 2.0), min_gain: float = 0.01, cache: Optional[str] = None, **opts) -> dict:
    """Picks the batched launch grids and workgroup sizes for a workload by measurement:
    coordinate descent over the families (default all grid families, then the shape families
    tile_lanes and band_lanes), trying each grid maximum times `factors` (rounded to a
    multiple of 8, within [8, 16384]) and each other compiled workgroup size of a shape family
    on a fresh fixed batch of `fluids` (`warmup` untimed
    and `steps` timed steps each) and keeping a value only when it is at least `min_gain`
    faster than the best so far.  One untimed trial with the current grids comes first (the
    process's first batch runs slower).  Leaves the best grids set, writes them to `cache`
    (JSON) when given, and returns {"grids", "shapes", "ms_per_step", "baseline_ms_per_step",
    "trials"}.  Results of a batch do not depend on the grids or sizes, only its speed."""
    fams = tuple(families) if families is not None else GRID_FAMILIES + tuple(SHAPE_FAMILIES)
    for name in fams:
        if name not in GRID_FAMILIES and name not in SHAPE_FAMILIES:
            raise KeyError(f"unknown batch grid or shape family {name!r}")
    best, shapes = batch_grids(), batch_shapes()
    _time_batch(fluids, fps, 1, warmup, opts)
    base = best_t = _time_batch(fluids, fps, steps, warmup, opts)
    trials = [{"grids": dict(best), "shapes": dict(shapes), "ms_per_step": base * 1e3}]
    for name in fams:
        if name in SHAPE_FAMILIES:  # the other compiled workgroup sizes
            for v in SHAPE_FAMILIES[name]:
                if v == shapes[name]:
                    continue
                set_batch_shapes({name: v})
                t = _time_batch(fluids, fps, steps, warmup, opts)
                trials.append({"family": name, "lanes": v, "ms_per_step": t * 1e3})
                if t < best_t * (1.0 - min_gain):
                    best_t, shapes[name] = t, v
                set_batch_shapes({name: shapes[name]})
            continue
        start = best[name]
        for f in factors:
            v = min(16384, max(8, int(round(start * f / 8.0)) * 8))
            if v == best[name]:
                continue
            set_batch_grids({name: v})
            t = _time_batch(fluids, fps, steps, warmup, opts)
            trials.append({"family": name, "workgroups": v, "ms_per_step": t * 1e3})
            if t < best_t * (1.0 - min_gain):
                best_t, best[name] = t, v
            set_batch_grids({name: best[name]})
    result = {"grids": best, "shapes": shapes, "ms_per_step": best_t * 1e3, "baseline_ms_per_step": base * 1e3,
              "trials": trials}
    if cache:
        save_batch_grids(cache, best, shapes, ms_per_step=best_t * 1e3, baseline_ms_per_step=base * 1e3,
                         members=len(fluids), steps=steps, trials=trials)
    return result
