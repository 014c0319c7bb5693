"""MI355X-native FLIP liquid step for pxlengine scenes (HIP on gfx950).

The compute lives in ``libpxlflip.so`` (sources in ``csrc/``, C ABI in
``include/pxlflip.h``); this package is its Python binding.
"""
from .flip import (FIELDS, SOLVERS, AgentGrid, Batch, DeviceBuffer, FluidError, FluidState, ImportedBuffer,  # noqa: F401
                   Server, check, load_library, make_opts, make_server_opts, seed_particles)

__all__ = ["FluidState", "FluidError", "Server", "Batch", "AgentGrid", "DeviceBuffer", "ImportedBuffer", "check",
           "seed_particles", "load_library", "make_opts", "make_server_opts", "FIELDS", "SOLVERS"]
