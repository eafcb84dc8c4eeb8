"""CPU checks of the C ABI: the library loads, exports every declared entry
point, and its host-side key parser follows the reference's wire rules
(dpf.rs:277-316) exactly like the oracle. No GPU compute is called."""
import os
import re
import subprocess

import numpy as np
import pytest

import morphogenesis_amd as mg
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "morphogen_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mg_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 20
    nm = subprocess.run(["nm", "-D", "--defined-only", mg.LIB_PATH], capture_output=True,
                        text=True, check=True).stdout
    exported = set(re.findall(r"\bT (mg_\w+)", nm))
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    for s in syms:
        assert getattr(mg.lib, s) is not None


def test_abi_version():
    assert mg.lib.mg_abi_version() == 1


def test_library_carries_gfx950_code_object():
    # The offload bundle embedded in our own build names its target.
    blob = open(mg.LIB_PATH, "rb").read()
    assert b"hipv4-amdgcn-amd-amdhsa--gfx950" in blob


@pytest.mark.parametrize("d", [0, 1, 8, 20, 27, 31, 32])
def test_key_parse_matches_oracle(d):
    if d == 0:
        key = bytes(35)
    else:
        key = oracle.keygen_bytes(d, (1 << d) - 1, seed=d)[1]
    info = mg.parse_key(key)
    ok = oracle.key_from_bytes(key)
    assert (info.domain_bits, info.party, info.root_t) == (ok.domain_bits, ok.party, ok.root_t)
    assert info.wire_len == len(key) == 19 + 18 * d + 16


def test_key_parse_errors_match_oracle():
    good = oracle.keygen_bytes(10, 5, seed=2)[0]
    cases = [bytes(0), bytes(34), good[:-1], good + b"\0"]
    b = bytearray(19 + 18 * 33 + 16)
    b[18] = 33
    cases.append(bytes(b))
    for c in cases:
        with pytest.raises(oracle.OracleError) as oe:
            oracle.key_from_bytes(c)
        with pytest.raises(mg.MgError) as me:
            mg.parse_key(c)
        expect = {oracle.ERR_KEY_FORMAT: mg.MG_ERR_KEY_FORMAT,
                  oracle.ERR_DOMAIN: mg.MG_ERR_DOMAIN}[oe.value.code]
        assert me.value.code == expect
        assert mg.lib.mg_last_error()


def test_device_open_without_gpu_reports_no_device():
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK):
        pytest.skip("a GPU may be present")
    with pytest.raises(mg.NoDevice):
        mg.GpuScanner(0)


def test_null_arguments_are_rejected():
    assert mg.lib.mg_scan(None, None, None, None, 1, None, None, None) == mg.MG_ERR_INVALID_ARG
    assert mg.lib.mg_matrix_upload(None, 0, None, 0) == mg.MG_ERR_INVALID_ARG
    assert mg.lib.mg_device_close(None) == mg.MG_OK


@pytest.mark.parametrize("d,target,seed", [(1, 1, 0), (8, 42, 1), (20, 123_456, 2),
                                           (27, (1 << 27) - 1, 3), (31, 1 << 30, 4), (32, 7, 5)])
def test_keygen_matches_oracle(d, target, seed):
    assert mg.generate_keys(d, target, seed) == oracle.keygen_bytes(d, target, seed)


def test_keygen_rejects_bad_arguments():
    with pytest.raises(mg.InvalidDomainBits):
        mg.generate_keys(0, 0, 0)
    with pytest.raises(mg.InvalidDomainBits):
        mg.generate_keys(33, 0, 0)
    with pytest.raises(mg.RowsError):
        mg.generate_keys(8, 256, 0)


def test_host_seeded_rows_match_oracle():
    from morphogenesis_amd.data import seeded_rows
    assert np.array_equal(seeded_rows(77, 0, 100, 48), oracle.fill_rows(77, 0, 100, 48))
    assert np.array_equal(seeded_rows(5, 123_456, 3, 32), oracle.fill_rows(5, 123_456, 3, 32))
