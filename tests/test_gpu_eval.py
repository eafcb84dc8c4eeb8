"""GPU tests of key evaluation: ChaChaKey::{eval, eval_subtree, full_eval}
(crates/morphogen-gpu-dpf/src/dpf.rs:103-246) through mg_key_eval /
mg_key_eval_subtree / mg_key_full_eval, against the oracle's restatement and
the DPF invariants the reference tests (dpf.rs:474-608: two-party XOR is the
point function, full_eval == point eval, subtree == slice of full_eval,
argument errors)."""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mg():
    import morphogenesis_amd
    return morphogenesis_amd


@pytest.fixture(scope="module")
def scanner(mg):
    s = mg.GpuScanner(0)
    yield s
    s.close()


def as_bytes(words):
    """Oracle leaves (uint32 words) in the wire byte order (little-endian)."""
    return np.ascontiguousarray(words, dtype="<u4").view(np.uint8).reshape(-1, 16)


@pytest.mark.parametrize("d", [1, 2, 5, 12, 20])
def test_full_eval_equals_oracle_and_is_a_point_function(mg, scanner, d):
    target = (1 << d) - 1 if d < 5 else (1 << d) // 3
    k0b, k1b = oracle.keygen_bytes(d, target, seed=d)
    f0 = scanner.full_eval(k0b)
    f1 = scanner.full_eval(k1b)
    assert f0.shape == (1 << d, 16)
    assert np.array_equal(f0, as_bytes(oracle.full_eval(oracle.key_from_bytes(k0b))))
    assert np.array_equal(f1, as_bytes(oracle.full_eval(oracle.key_from_bytes(k1b))))
    x = f0 ^ f1
    assert np.all(x[target] == 0xFF)
    assert not np.any(np.delete(x, target, axis=0))


def test_point_eval_equals_full_eval_and_ignores_high_bits(mg, scanner):
    d = 16
    k0b, _ = oracle.keygen_bytes(d, 12345, seed=4)
    full = scanner.full_eval(k0b)
    rng = np.random.default_rng(0)
    idx = rng.integers(0, 1 << d, size=5000).astype(np.uint64)
    pts = scanner.eval(k0b, idx)
    assert np.array_equal(pts, full[idx.astype(np.int64)])
    # Only the low d bits of an index are read (dpf.rs:111).
    hi = idx + (np.uint64(1) << np.uint64(d)) * np.uint64(5)
    assert np.array_equal(scanner.eval(k0b, hi), pts)
    ok = oracle.key_from_bytes(k0b)
    for i in (0, 1, 12345, (1 << d) - 1):
        assert np.array_equal(scanner.eval(k0b, [i])[0], as_bytes(oracle.eval_point(ok, i))[0])
    assert scanner.eval(k0b, []).shape == (0, 16)


def test_subtrees_are_slices_of_full_eval(mg, scanner):
    d = 14
    k0b, _ = oracle.keygen_bytes(d, 777, seed=5)
    full = scanner.full_eval(k0b)
    ok = oracle.key_from_bytes(k0b)
    for start, size in ((0, 1), (777, 1), (512, 256), (4096, 4096), (8192, 8192), (0, 1 << d),
                        (64, 64), (16380, 4)):
        sub = scanner.eval_subtree(k0b, start, size)
        assert np.array_equal(sub, full[start:start + size]), (start, size)
        assert np.array_equal(sub, as_bytes(oracle.eval_subtree(ok, start, size))), (start, size)


def test_argument_errors(mg, scanner):
    k0b, _ = oracle.keygen_bytes(10, 3, seed=6)
    with pytest.raises(mg.InvalidKeyFormat):
        scanner.eval_subtree(k0b, 0, 3)        # size not a power of two
    with pytest.raises(mg.InvalidKeyFormat):
        scanner.eval_subtree(k0b, 8, 16)       # start not aligned to the size
    with pytest.raises(mg.RowsError):
        scanner.eval_subtree(k0b, 0, 2048)     # larger than the 2^10-leaf domain
    with pytest.raises(mg.InvalidKeyFormat):
        scanner.eval(k0b[:-1], [0])            # wire length rule
    out = np.zeros((1 << 10) - 1, dtype=np.uint8)
    kb = np.frombuffer(k0b, dtype=np.uint8).copy()
    rc = mg.lib.mg_key_full_eval(scanner._h, kb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                 len(k0b), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                 out.nbytes)
    assert rc == mg.MG_ERR_INVALID_ARG         # InvalidKeyLength


def test_large_domains_through_chunks(mg, scanner):
    # d = 25: 2^25 leaves (512 MiB) in two 2^24-leaf chunks.
    d, target = 25, 23_456_789
    k0b, k1b = oracle.keygen_bytes(d, target, seed=7)
    f0 = scanner.full_eval(k0b)
    f1 = scanner.full_eval(k1b)
    x = f0 ^ f1
    assert np.all(x[target] == 0xFF)
    x[target] = 0
    assert not np.any(x)
    del x, f1
    ok = oracle.key_from_bytes(k0b)
    rng = np.random.default_rng(8)
    for i in rng.integers(0, 1 << d, size=200):
        assert np.array_equal(f0[i], as_bytes(oracle.eval_point(ok, int(i)))[0]), i
    # d = 31: the last 2^24-leaf subtree of the domain, spot-checked.
    d = 31
    k0b, k1b = oracle.keygen_bytes(d, (1 << 31) - 5, seed=9)
    start, size = (1 << 31) - (1 << 24), 1 << 24
    s0 = scanner.eval_subtree(k0b, start, size)
    s1 = scanner.eval_subtree(k1b, start, size)
    x = s0 ^ s1
    assert np.all(x[size - 5] == 0xFF)
    x[size - 5] = 0
    assert not np.any(x)
    ok = oracle.key_from_bytes(k0b)
    for i in (0, 1, 12345, size - 5, size - 1):
        assert np.array_equal(s0[i], as_bytes(oracle.eval_point(ok, start + i))[0]), i
