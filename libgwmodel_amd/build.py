"""Build recipe of the gfx950 HIP library: libgwmodel_amd/lib/libgwmodel_hip.so.

Every translation unit is compiled by hipcc for gfx950 only, with
-ffp-contract=off (no implicit FMA: the distance and weight expressions keep
the reference's roundings) and packed FP32 arithmetic disabled.  Objects are
built in parallel and linked into one shared library exporting the C ABI of
include/gwmodel_hip.h.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(OUT_DIR, "obj")
LIB = os.path.join(OUT_DIR, "libgwmodel_hip.so")
FACADE_LIB = os.path.join(OUT_DIR, "libgwmodelpp_hip.so")
FACADE_SRC = os.path.join(HERE, "cpp", "gwr_facade.cpp")
FACADE_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_facade.cpp")
FACADE_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_facade")
TASK_SRC = os.path.join(HERE, "cpp", "gwr_gpu_task.cpp")
TASK_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_gpu_task.cpp")
TASK_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_gpu_task")

SOURCES = ["gwm_contract.hip", "gwm_contract_mixed.hip", "gwm_select.hip", "gwm_solve.hip", "gwm_varsel.hip",
           "gwm_mgwr_hat.hip", "gwm_gemm.hip", "gwm_plan.hip", "gwm_noalloc.hip", "gwm_problem.cpp", "gwm_sweep.cpp",
           "gwm_api.cpp", "gwm_batch.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"),
]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [HIPCC] + FLAGS + lang + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + [
        "-lpthread", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    build_facade()
    if verbose:
        print("built", LIB, FACADE_LIB, FACADE_TEST_BIN)
    return LIB


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("%s\n%s" % (" ".join(cmd), r.stderr))


def build_facade():
    """The C++ host interface (include/gwmodelpp/gwmodel.h) and the
    IGWRBasicGpuTask entry (include/gwmodelpp/IGWRBasicGpuTask.h) are host-only
    C++ over the C ABI: g++, linked against libgwmodel_hip.so, plus their KAT
    programs tests/cpp/test_facade and tests/cpp/test_gpu_task."""
    inc = "-I" + os.path.join(ROOT, "include")
    tmp = FACADE_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", inc, FACADE_SRC, TASK_SRC, "-o", tmp, "-L" + OUT_DIR,
          "-lgwmodel_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FACADE_LIB)
    os.makedirs(os.path.dirname(FACADE_TEST_BIN), exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", inc, FACADE_TEST_SRC, "-o", FACADE_TEST_BIN, "-L" + OUT_DIR,
          "-lgwmodelpp_hip", "-lgwmodel_hip", "-Wl,-rpath,$ORIGIN/../../../libgwmodel_amd/lib"])
    _run(["g++", "-O2", "-std=c++17", inc, TASK_TEST_SRC, "-o", TASK_TEST_BIN, "-L" + OUT_DIR,
          "-lgwmodelpp_hip", "-lgwmodel_hip", "-Wl,-rpath,$ORIGIN/../../../libgwmodel_amd/lib"])


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
