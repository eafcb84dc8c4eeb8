"""mg_scan_noalloc against mg_scan: the same keys and options give the same
answers when every per-call device buffer sits in the library's fixed scratch
array. Two calls in a row on each shape, the second smaller, so the second
call runs on whatever the first left in the array; each answer is also
checked against the oracle."""
import ctypes

import numpy as np
import pytest

import morphogenesis_amd as mg
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scanner():
    s = mg.GpuScanner(0)
    yield s
    s.close()


def keys_for(d, n_rows, n, seed):
    out = []
    for i in range(n):
        t = (seed * 7919 + i * 104729) % n_rows
        out.append(mg.generate_keys(d, t, seed=seed * 1000 + i)[i % 2])
    return out


def check_pair_of_calls(scanner, shapes, opts_fn=None):
    """shapes: [(n_rows, row_bytes, d, n_keys, seed), ...] run back to back."""
    for n_rows, w, d, nk, seed in shapes:
        print(f"noalloc case: {n_rows} x {w} B, d={d}, {nk} keys, seed {seed}", flush=True)
        m = mg.GpuPageMatrix(scanner, n_rows, w)
        m.fill_seeded(seed)
        keys = keys_for(d, n_rows, nk, seed)
        opts = opts_fn(scanner, m) if opts_fn else None
        a_free, t = scanner.scan_keys_noalloc(m, keys, opts, with_timing=True)
        a_ref = scanner.scan_keys(m, keys, opts)
        assert np.array_equal(a_free, a_ref), (n_rows, w, nk)
        rows = oracle.fill_rows(seed, 0, n_rows, w)
        if opts_fn is not None and getattr(opts, "delta", None):
            pass  # delta cases are compared with mg_scan only (checked against the oracle there)
        else:
            assert np.array_equal(a_free, oracle.scan_tiled(keys, rows, 8))
        assert t["scan_kernel_ms"] > 0
        m.close()


def test_lane_dfs_two_calls(scanner):
    check_pair_of_calls(scanner, [(1 << 18, 32, 18, 12, 3), (5000, 32, 13, 3, 4)])


def test_split_two_calls(scanner):
    check_pair_of_calls(scanner, [(40000, 1024, 16, 4, 5), (9000, 1024, 14, 1, 6)])


def test_stream_two_calls(scanner):
    check_pair_of_calls(scanner, [(30000, 272, 15, 3, 7), (1000, 144, 10, 2, 8)])


def test_with_delta_two_calls(scanner):
    deltas = []

    def with_delta(sc, m):
        rng = np.random.default_rng(m.n_rows)
        n = 300
        idx = rng.integers(0, m.n_rows, size=n).astype(np.uint64)
        diffs = rng.integers(0, 256, size=(n, m.row_bytes), dtype=np.uint8)
        dl = mg.DeltaBuffer(sc, m.row_bytes, n)
        dl.set(idx, diffs, epoch=1)
        deltas.append(dl)
        return mg.make_opts(delta=dl)

    check_pair_of_calls(scanner, [(1 << 16, 32, 16, 6, 9), (4000, 32, 12, 2, 10)], with_delta)
    for dl in deltas:
        dl.close()


def test_refusals_are_host_side(scanner):
    m = mg.GpuPageMatrix(scanner, 1024, 32)
    m.fill_seeded(1)
    key = mg.generate_keys(10, 5, seed=1)[0]
    limit = 1024  # MG_NOALLOC_MAX_KEYS
    with pytest.raises(mg.MgError) as e:
        scanner.scan_keys_noalloc(m, [key] * (limit + 1))
    assert e.value.code == 1 and "no-alloc limit" in str(e.value)
    with pytest.raises(mg.MgError) as e:
        scanner.scan_keys_noalloc(m, [key], mg.make_opts(use_graph=True))
    assert e.value.code == 1
    # The device is still usable, and an empty call is a no-op.
    assert scanner.scan_keys_noalloc(m, []).shape == (0, 32)
    ans = scanner.scan_keys_noalloc(m, [key])
    assert np.array_equal(ans, scanner.scan_keys(m, [key]))
    m.close()
