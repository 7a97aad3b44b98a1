"""Runs the C++ mirror's known-answer program (tests/cpp/test_api.cpp) on
the GPU: every hot-path test of the reference, through the C ABI."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(__file__), "cpp", "test_api")


def test_cpp_api_reference_kats():
    assert os.path.exists(EXE), "tests/cpp/test_api not built (run __graft_entry__.build())"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("PASS") >= 29
