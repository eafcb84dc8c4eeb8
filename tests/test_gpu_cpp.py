"""The compiled C++ host-interface tests on the GPU: the reference's kernel
tests (crates/morphogen-gpu-dpf/src/kernel.rs:1443-1690) restated against
include/morphogen_hip.hpp, each answer also checked against the oracle. The
binary runs as a child process of its own."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_scan_cpp")


def test_cpp_interface_on_the_gpu():
    assert os.path.exists(BIN), "build the tests first (__graft_entry__.build())"
    print("running", BIN, "--gpu", flush=True)
    r = subprocess.run([BIN, "--gpu"], capture_output=True, text=True, timeout=240)
    print(r.stdout[-4000:], r.stderr[-4000:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout
