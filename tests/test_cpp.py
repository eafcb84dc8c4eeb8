"""The compiled C++ host-interface tests (tests/cpp/test_scan.cpp) in their
no-GPU mode: parameter and key errors, keygen == oracle keygen, and the
scanner reporting a missing device instead of falling back."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_scan_cpp")


@pytest.fixture(scope="module")
def binary():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-C", os.path.dirname(BIN)], check=True)
    return BIN


def test_cpp_interface_without_a_gpu(binary):
    r = subprocess.run([binary, "--cpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout
