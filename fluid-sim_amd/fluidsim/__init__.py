"""fluidsim: MI355X-native stable-fluids step behind the C ABI of include/fluidsim.h.

The compute path is libfluidsim.so (hand-written HIP for gfx950); this package
is the host-side mirror of the reference's src/gpu solver interface.
"""
from ._lib import FluidSimError, NotConverged, library_path, load
from .solver import GPUFluidSolver, createDevice
from . import initial
from .batch import CavityBatch, random_cavities

__all__ = ["GPUFluidSolver", "createDevice", "FluidSimError", "NotConverged", "initial", "load",
           "library_path"]
