"""The C++ host interfaces on the device: tests/cpp/test_facade restates the
reference's Catch2 KAT sections (test/testGWRBasic.cpp:40-240,
test/testGWRMultiscale.cpp:45-88, 135-173) on include/gwmodelpp/gwmodel.h with
ParallelType::HIP; tests/cpp/test_gpu_task drives the IGWRBasicGpuTask entry
(include/gwmodelpp/IGWRBasicGpuTask.h) through its factory functions."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_facade")
TASK_BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_gpu_task")

pytestmark = pytest.mark.gpu


def test_cpp_facade_kats():
    assert os.path.exists(BIN), "build the library first (__graft_entry__.build())"
    r = subprocess.run([BIN, os.path.join(ROOT, "tests", "golden")], capture_output=True, text=True, timeout=150)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "FACADE TESTS PASSED" in r.stdout


def test_cpp_gpu_task_kats():
    assert os.path.exists(TASK_BIN), "build the library first (__graft_entry__.build())"
    r = subprocess.run([TASK_BIN, os.path.join(ROOT, "tests", "golden")], capture_output=True, text=True,
                       timeout=150)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "GPU TASK TESTS PASSED" in r.stdout
