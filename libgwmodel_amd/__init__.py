"""MI355X (gfx950) back-end for libgwmodel's GWR hot path.

The compute runs in hand-written HIP kernels behind the C ABI of
include/gwmodel_hip.h (libgwmodel_amd/lib/libgwmodel_hip.so); this package
holds the ctypes binding (capi), the synthetic workload generators (synth)
and the build recipe (build).
"""
from . import capi  # noqa: F401
from .capi import Context, GwmError  # noqa: F401
