"""The wire parser/formatter and client functions under the address and
undefined-behaviour sanitizers (tests/cpp/fuzz_wire.cpp): mutated and random
request bodies, deep nesting, oversized strings, short output buffers. The
harness is host code only and runs without a GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(CPP, "fuzz_wire_asan")


@pytest.fixture(scope="module")
def harness():
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    subprocess.run(["make", "-s", "-C", CPP, "fuzz_wire_asan"], check=True)
    return BIN


@pytest.mark.parametrize("seed", [11, 12])
def test_wire_and_client_under_sanitizers(harness, seed):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([harness, "20000", str(seed)], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "0 failure(s)" in r.stdout
