"""The C++ host interface (include/gwmodelpp/gwmodel.h) on the device: runs
tests/cpp/test_facade, which restates the reference's Catch2 KAT sections
(test/testGWRBasic.cpp:40-137, test/testGWRMultiscale.cpp:135-173) with
ParallelType::HIP."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_facade")

pytestmark = pytest.mark.gpu


def test_cpp_facade_kats():
    assert os.path.exists(BIN), "build the library first (__graft_entry__.build())"
    r = subprocess.run([BIN, os.path.join(ROOT, "tests", "golden")], capture_output=True, text=True, timeout=150)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "FACADE TESTS PASSED" in r.stdout
