"""MI355X-native lattice-Boltzmann + k-omega time step of kOmegaFoulingModel.

The compute path is libkofm.so (HIP kernels for gfx950 behind the C ABI in
include/kofm.h); ``lib.Lattice`` is its ctypes binding and ``channel`` builds
wall descriptions and the named workloads.
"""
from . import channel, lib  # noqa: F401
from .lib import Lattice, KofmError, KofmLibraryMissing  # noqa: F401
