"""MI355X-native FLIP liquid step for pxlengine scenes (HIP on gfx950).

The compute lives in ``libpxlflip.so`` (sources in ``csrc/``, C ABI in
``include/pxlflip.h``); this package is its Python binding.
"""
from .flip import (FIELDS, GRID_FAMILIES, SOLVERS, AgentGrid, Batch, DeviceBuffer, FluidError,  # noqa: F401
                   FluidState, ImportedBuffer, ImportedMemory, SHAPE_FAMILIES, Server, autotune_batch, batch_grids,
                   batch_shapes, check,
                   check_scene, load_batch_grids, load_library, make_opts, make_server_opts, save_batch_grids,
                   seed_particles, set_batch_grids, set_batch_shapes)

__all__ = ["FluidState", "FluidError", "Server", "Batch", "AgentGrid", "DeviceBuffer", "ImportedBuffer",
           "ImportedMemory", "check", "check_scene", "seed_particles", "load_library", "make_opts",
           "make_server_opts", "FIELDS", "SOLVERS", "GRID_FAMILIES", "batch_grids", "set_batch_grids",
           "save_batch_grids", "load_batch_grids", "autotune_batch", "SHAPE_FAMILIES", "batch_shapes",
           "set_batch_shapes"]
