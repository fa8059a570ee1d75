"""onionpirv2_amd -- MI355X-native OnionPIR v2 server answer path.

The server path runs as hand-written HIP kernels for gfx950 behind the C ABI
of include/onionpir_gpu.h (libonionpir.so, built in-tree by
``python -m onionpirv2_amd.build``).  ``GpuServer`` is the thin host handle.
"""
from ._capi import OnionError, OnionKeyError, OnionValueError, lib  # noqa: F401
from .server import GpuServer  # noqa: F401
