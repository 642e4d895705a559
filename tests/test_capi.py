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
