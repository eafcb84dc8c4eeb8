"""GPU tests of matrix ingestion from raw row files and snapshot hot-swap.

Files follow the reference's loader (bench_page_gpu_batch.rs:326-379: ragged
tail zero-padded) and snapshot decode rule (snapshot_config.rs:144-159: empty
or ragged refused); rows in the 32-byte Compact schema
(docs/DATA_STRUCTURES.md:17-26) decode back through AccountData after a
two-party query. Snapshot swaps follow EpochManager::submit_snapshot
(epoch.rs:407-463): requests before it answer for the old rows, requests
after it for the new rows at the new epoch, pending updates dropped.
"""
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope="module")
def mg():
    import morphogenesis_amd
    return morphogenesis_amd


@pytest.fixture(scope="module")
def scanner(mg):
    s = mg.GpuScanner(0)
    yield s
    s.close()


def write(path, data: bytes):
    with open(path, "wb") as f:
        f.write(data)
    return str(path)


def test_compact_file_round_trip_and_query(mg, scanner, tmp_path):
    from morphogenesis_amd.data import AccountData, compact_rows
    n = 1 << 14
    rng = np.random.default_rng(0)
    bal = rng.integers(0, 2**63, size=n, dtype=np.uint64)
    non = rng.integers(0, 2**40, size=n, dtype=np.uint64)
    cid = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
    rows = compact_rows(bal, non, cid)
    path = write(tmp_path / "accounts.bin", rows.tobytes())
    m = mg.GpuPageMatrix.from_file(scanner, path, 32, pad_tail=False)
    try:
        assert (m.n_rows, m.row_bytes) == (n, 32)
        assert np.array_equal(m.download(), rows)
        for t in (0, 777, n - 1):
            k0, k1 = mg.generate_keys(14, t, seed=t + 1)
            a0, a1 = scanner.scan_keys(m, [k0, k1])
            acct = AccountData.from_bytes((a0 ^ a1).tobytes())
            assert (acct.balance, acct.nonce, acct.code_id) == (int(bal[t]), int(non[t]),
                                                                int(cid[t]))
        keys = [oracle.keygen_bytes(14, t, seed=9)[t % 2] for t in (3, 5000, 16000)]
        assert np.array_equal(scanner.scan_keys(m, keys), oracle.scan_tiled(keys, rows, THREADS))
    finally:
        m.close()


def test_ragged_empty_and_bad_files(mg, scanner, tmp_path):
    rng = np.random.default_rng(1)
    data = rng.integers(0, 256, size=32 * 1000 + 7, dtype=np.uint8).tobytes()
    path = write(tmp_path / "ragged.bin", data)
    m = mg.GpuPageMatrix.from_file(scanner, path, 32, pad_tail=True)
    try:
        assert m.n_rows == 1001
        got = m.download().tobytes()
        assert got[:len(data)] == data and got[len(data):] == bytes(32 - 7)
    finally:
        m.close()
    with pytest.raises(mg.RowsError):
        mg.GpuPageMatrix.from_file(scanner, path, 32, pad_tail=False)
    with pytest.raises(mg.RowsError):
        mg.GpuPageMatrix.from_file(scanner, write(tmp_path / "empty.bin", b""), 32)
    with pytest.raises(mg.RowsError):
        mg.GpuPageMatrix.from_file(scanner, path, 24)
    with pytest.raises(mg.MgError):
        mg.GpuPageMatrix.from_file(scanner, str(tmp_path / "missing.bin"), 32)


def test_file_spanning_many_staging_buffers(mg, scanner, tmp_path):
    # 150 MiB + 16 bytes: three 64 MiB staging rounds and a ragged 1 KiB-row tail.
    w = 1024
    size = 150 * (1 << 20) + 16
    data = np.random.default_rng(2).integers(0, 256, size=size, dtype=np.uint8)
    path = write(tmp_path / "big.bin", data.tobytes())
    m = mg.GpuPageMatrix.from_file(scanner, path, w)
    try:
        assert m.n_rows == (size + w - 1) // w
        got = m.download().reshape(-1)
        assert np.array_equal(got[:size], data) and not got[size:].any()
    finally:
        m.close()
    os.unlink(path)


@pytest.mark.parametrize("double_buffer", [False, True])
def test_snapshot_swap(mg, scanner, tmp_path, double_buffer):
    d, w = 14, 32
    rows_a = oracle.fill_rows(81, 0, 1 << d, w)
    rows_b = oracle.fill_rows(82, 0, 10000, w)  # fewer rows, same domain
    a = mg.GpuPageMatrix.new(scanner, rows_a)
    b = mg.GpuPageMatrix.from_file(scanner, write(tmp_path / "b.bin", rows_b.tobytes()), w,
                                   pad_tail=False)
    srv = mg.GpuServer(scanner, a, segments=4, lane_depth=4, double_buffer=double_buffer)
    try:
        keys1 = [oracle.keygen_bytes(d, t, seed=100 + t)[0] for t in (1, 2, 9000)]
        keys2 = [oracle.keygen_bytes(d, t, seed=200 + t)[1] for t in (4, 5, 9999)]
        # An update the snapshot will drop (the new rows already hold the truth).
        idx = np.array([1, 2], dtype=np.uint64)
        diffs = np.full((2, w), 0x5A, dtype=np.uint8)
        srv.update(idx, diffs, 3)
        state_a = rows_a.copy()
        state_a[1] ^= diffs[0]
        state_a[2] ^= diffs[1]
        t1 = srv.submit(keys1)
        srv.submit_snapshot(b, 10)
        ans1, info1 = srv.wait(t1, 3)
        assert info1["epoch"] == 3
        assert np.array_equal(ans1, oracle.scan_tiled(keys1, state_a, THREADS))
        t2 = srv.submit(keys2)
        ans2, info2 = srv.wait(t2, 3)
        assert info2["epoch"] == 10
        assert np.array_equal(ans2, oracle.scan_tiled(keys2, rows_b, THREADS))
        st = srv.stats()
        assert st["snapshots"] == 1 and st["n_rows"] == 10000 and st["delta_entries"] == 0
        # Updates now address the new, shorter matrix.
        with pytest.raises(mg.RowsError):
            srv.update(np.array([12000], dtype=np.uint64), np.zeros((1, w), np.uint8), 11)
        srv.update(np.array([9999], dtype=np.uint64), np.full((1, w), 7, np.uint8), 11)
        state_b = rows_b.copy()
        state_b[9999] ^= 7
        t3 = srv.submit(keys2)
        ans3, info3 = srv.wait(t3, 3)
        assert info3["epoch"] == 11
        assert np.array_equal(ans3, oracle.scan_tiled(keys2, state_b, THREADS))
        # The old matrix is no longer read; it can go.
        a.close()
        # Bad snapshots: another row width, or more rows than the domain.
        c = mg.GpuPageMatrix(scanner, 100, 64)
        with pytest.raises(mg.RowsError):
            srv.submit_snapshot(c, 12)
        c.close()
        big = mg.GpuPageMatrix(scanner, (1 << d) + 1, w)
        with pytest.raises(mg.RowsError):
            srv.submit_snapshot(big, 12)
        big.close()
    finally:
        srv.close()
        b.close()
