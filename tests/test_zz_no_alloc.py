"""gwm_hip_criterion_noalloc (the allocation-free entry point) against
gwm_hip_criterion on the same data, through the C ABI.

The inputs live in device arrays this test allocates with the HIP runtime the
library itself links (/opt/rocm/lib/libamdhip64.so.7); the entry point keeps
everything else in its static per-device scratch array.  Tolerance: the two
entry points sum in different orders (spatially sorted data and 16 segments
against the caller's order and at most 4), so criteria agree within 1e-11
relative, the FP64 bar of the other criterion tests.
"""
import ctypes
import sys

import numpy as np
import pytest

from libgwmodel_amd import capi, synth

pytestmark = pytest.mark.gpu

_HIP = "/opt/rocm/lib/libamdhip64.so.7"


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class DeviceArrays:
    """Device copies of float64 host arrays (column-major), freed on close."""

    def __init__(self):
        self.hip = ctypes.CDLL(_HIP)
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def put(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0   # host to device
        return p.value

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


@pytest.fixture(scope="module")
def ctx():
    say("create context")
    c = capi.Context(0)
    yield c
    c.close()


def both(ctx, dev, d, distance, kernel, adaptive, bw, crit):
    """(noalloc status, noalloc value, gwm_hip_criterion value) on d."""
    n, k = d["X"].shape
    ctx.set_data(d["coords"], d["X"], d["y"], distance)
    ref = ctx.criterion(kernel, adaptive, bw, crit)
    rc, v = ctx.criterion_noalloc(n, k, dev.put(d["coords"]), dev.put(d["X"]), dev.put(d["y"]), kernel, adaptive, bw,
                                  crit, distance)
    return rc, v, ref


CASES = [
    ("uniform_square", capi.EUCLIDEAN, capi.BISQUARE, True, 60.0, capi.CRIT_CV),
    ("uniform_square", capi.EUCLIDEAN, capi.GAUSSIAN, False, 8.0, capi.CRIT_AIC),
    ("uniform_square", capi.EUCLIDEAN, capi.EXPONENTIAL, False, 3.0, capi.CRIT_CV),
    ("sphere_box", capi.GREAT_CIRCLE, capi.GAUSSIAN, False, 400.0, capi.CRIT_AIC),
    ("sphere_box", capi.GREAT_CIRCLE, capi.TRICUBE, True, 80.0, capi.CRIT_CV),
]


def data(name, n):
    return synth.uniform_square(n=n) if name == "uniform_square" else synth.sphere_box(n=n, k=6)


@pytest.mark.parametrize("name,distance,kernel,adaptive,bw,crit", CASES)
def test_noalloc_matches_criterion_twice(ctx, name, distance, kernel, adaptive, bw, crit):
    """Two calls in a row, the second on a smaller input: the scratch array
    keeps the first call's contents, which must not leak into the second."""
    dev = DeviceArrays()
    try:
        for n in (3000, 700):
            say("case: noalloc %s n=%d kernel=%d adaptive=%d crit=%d" % (name, n, kernel, adaptive, crit))
            rc, v, ref = both(ctx, dev, data(name, n), distance, kernel, adaptive, bw, crit)
            assert rc == capi.GWM_OK, ctx.last_error_message()
            assert ref < capi.DBL_MAX
            assert v == pytest.approx(ref, rel=1e-11), (n, v, ref)
    finally:
        dev.close()


def test_noalloc_not_spd_matches(ctx):
    say("case: noalloc not SPD")
    rng = np.random.default_rng(9)
    n = 300
    x1 = rng.normal(size=n)
    d = dict(coords=rng.uniform(0, 10, (n, 2)), X=np.column_stack([np.ones(n), x1, x1]), y=rng.normal(size=n))
    dev = DeviceArrays()
    try:
        rc, v, ref = both(ctx, dev, d, capi.EUCLIDEAN, capi.GAUSSIAN, False, 3.0, capi.CRIT_CV)
        assert ref == capi.DBL_MAX
        assert rc == capi.GWM_ERR_NOT_SPD and v == capi.DBL_MAX
    finally:
        dev.close()


def test_noalloc_rejects_too_large(ctx):
    """Sizes beyond the scratch array are refused before any device work."""
    say("case: noalloc too large")
    dev = DeviceArrays()
    try:
        n, k = 200_000, 8
        p = dev.put(np.zeros(4))   # never read: the size check comes first
        rc, v = ctx.criterion_noalloc(n, k, p, p, p, capi.GAUSSIAN, False, 1.0, capi.CRIT_AIC)
        assert rc == capi.GWM_ERR_INVALID and v == capi.DBL_MAX
        assert "MiB of scratch" in ctx.last_error_message()
    finally:
        dev.close()
