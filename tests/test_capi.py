"""CPU checks of the C-ABI library: it loads without a GPU and exports every
symbol that include/gwmodel_hip.h declares (no compute calls here)."""
import ctypes

from libgwmodel_amd import capi


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    declared = capi.header_symbols()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(lib, name), name
        assert name in capi.SIGNATURES, name


def test_device_count_without_gpu_is_safe():
    n = capi.device_count()
    assert n >= 0


def test_create_fails_cleanly_without_device():
    if capi.device_count() > 0:
        return
    h = ctypes.c_void_p()
    rc = capi.load().gwm_hip_create(0, ctypes.byref(h))
    assert rc == capi.GWM_ERR_NO_DEVICE


def test_gpu_task_library_exports_the_factories():
    """include/gwmodelpp/IGWRBasicGpuTask.h: the three extern "C" names of the
    reference's gwmodelcuda entry (IGWRBasicGpuTask.h:63-65)."""
    import os
    import re
    capi.load()   # the host library depends on the C-ABI library
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lib = ctypes.CDLL(os.path.join(root, "libgwmodel_amd", "lib", "libgwmodelpp_hip.so"))
    text = open(os.path.join(root, "include", "gwmodelpp", "IGWRBasicGpuTask.h")).read()
    names = sorted(set(re.findall(r'extern "C" GWMODELCUDA_API [^(]*?\b(\w+)\s*\(', text)))
    assert names == ["GWRBasicGpuTaskFit_Create", "GWRBasicGpuTaskPredict_Create", "GWRBasicGpuTask_Del"]
    for name in names:
        assert hasattr(lib, name), name


def test_sqrt_boundary_decides_the_distance_test():
    """gwm_hip_sqrt_boundary(bw) = T, the smallest double whose correctly
    rounded square root is >= bw: then sqrt(T) >= bw > sqrt(T-), so every
    squared distance d2 satisfies d2 < T exactly when sqrt(d2) < bw (the
    FP64 bisquare and boxcar weights decide BandwidthWeight.h's d < bw
    this way).  Checked on the defining property with Python's correctly
    rounded math.sqrt, over random bandwidths of every magnitude, squares of
    doubles (where sqrt(d^2) = d) and the subnormal / overflow ends."""
    import math
    import random
    lib = capi.load()
    out = ctypes.c_double(0.0)
    rng = random.Random(12)
    bws = [1.0, 2.0, 0.5, 3.0, 61.890021059082855, 1e-3, 1e5, 7.25, 2.0 ** 511, 2.0 ** 512, 1.3407807929942596e154,
           1.7976931348623157e308, 2.0 ** -500, 2.0 ** -500 * (1 - 2 ** -53), 1e-160, 1e-300, 5e-324, 2.2e-308]
    bws += [rng.uniform(0, 1000) for _ in range(3000)]
    bws += [math.ldexp(rng.random() + 0.5, rng.randint(-1070, 1020)) for _ in range(3000)]
    bws += [math.sqrt(rng.uniform(0, 1e6)) for _ in range(1000)]   # distances that are roots of doubles
    for bw in bws:
        assert lib.gwm_hip_sqrt_boundary(bw, ctypes.byref(out)) == 0
        T = out.value
        assert math.sqrt(T) >= bw, bw
        below = math.nextafter(T, 0.0)
        assert T == 0.0 or math.sqrt(below) < bw, bw
        # the squared distance of a point exactly at bw is not below T
        sq = bw * bw
        if 0.0 < sq < math.inf and sq >= 2.0 ** -1022:
            assert not (sq < T), bw
    for bw in (0.0, -1.0, -math.inf):
        assert lib.gwm_hip_sqrt_boundary(bw, ctypes.byref(out)) == 0 and out.value == 0.0
    assert lib.gwm_hip_sqrt_boundary(math.inf, ctypes.byref(out)) == 0 and out.value == math.inf
