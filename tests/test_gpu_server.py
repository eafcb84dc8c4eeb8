"""GPU tests of the continuous-batching request loop (mg_server_*).

Each result must be the same alone and in the mix: a request's answers equal
a plain whole-matrix scan (and the oracle) of the matrix state at its
admission, whatever requests and updates are in flight with it. Updates follow
the reference's epoch semantics (an update applies to requests after it;
scan.rs:260-299, epoch.rs:465-549), including merges into a base copy.
"""
import os
import threading
import time

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


def make_keys(d, targets, party, seed):
    return [oracle.keygen_bytes(d, t, seed=seed + i)[party] for i, t in enumerate(targets)]


def test_request_alone_equals_request_in_mix(mg, scanner):
    n, w, d = 1 << 18, 32, 18
    rows = oracle.fill_rows(61, 0, n, w)
    m = mg.GpuPageMatrix.new(scanner, rows)
    srv = mg.GpuServer(scanner, m, segments=8, lane_depth=4)
    assert srv.stats()["segments"] == 8
    rng = np.random.default_rng(0)
    reqs = []
    for r in range(12):
        k = 3 * int(rng.integers(1, 5))  # 3 cuckoo keys x k lookups
        reqs.append(make_keys(d, [int(x) for x in rng.integers(0, n, size=k)], r % 2, 100 * r))
    alone = []
    for keys in reqs[:3]:
        t = srv.submit(keys)
        ans, info = srv.wait(t, len(keys))
        assert info["n_keys"] == len(keys) and info["latency_ms"] >= info["service_ms"]
        alone.append(ans)
    # The same requests again, interleaved with others arriving while they run.
    tickets = []
    for i, keys in enumerate(reqs):
        tickets.append((i, srv.submit(keys)))
        if i % 4 == 3:
            time.sleep(0.002)
    mixed = {i: srv.wait(t, len(reqs[i]))[0] for i, t in tickets}
    for i in range(3):
        assert np.array_equal(mixed[i], alone[i])
    for i, keys in enumerate(reqs):
        direct = scanner.scan_keys(m, keys)
        assert np.array_equal(mixed[i], direct), i
    for i in (0, 5, 11):
        assert np.array_equal(mixed[i], oracle.scan_tiled(reqs[i], rows, THREADS))
    st = srv.stats()
    assert st["completed"] == 15 and st["active_keys"] == 0 and st["queued_requests"] == 0
    srv.close()
    m.close()


def test_concurrent_clients(mg, scanner):
    n, w, d = 1 << 16, 32, 16
    rows = oracle.fill_rows(62, 0, n, w)
    m = mg.GpuPageMatrix.new(scanner, rows)
    srv = mg.GpuServer(scanner, m, segments=4, lane_depth=4, max_active_keys=24)
    results, errors = {}, []

    def client(cid):
        try:
            rng = np.random.default_rng(cid)
            for j in range(4):
                keys = make_keys(d, [int(x) for x in rng.integers(0, n, size=6)], 0,
                                 1000 * cid + 10 * j)
                t = srv.submit(keys)
                results[(cid, j)] = (keys, srv.wait(t, len(keys))[0])
        except Exception as e:  # surfaced by the assert below
            errors.append(e)

    th = [threading.Thread(target=client, args=(c,)) for c in range(6)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errors, errors
    assert len(results) == 24
    for (cid, j), (keys, ans) in results.items():
        assert np.array_equal(ans, scanner.scan_keys(m, keys)), (cid, j)
    srv.close()
    m.close()


@pytest.mark.parametrize("double_buffer", [False, True])
def test_updates_pin_epoch_at_admission(mg, scanner, double_buffer):
    n, w, d = 1 << 16, 32, 16
    rows = oracle.fill_rows(63, 0, n, w)
    m = mg.GpuPageMatrix.new(scanner, rows)
    # A small delta capacity forces merges into a base copy along the way.
    srv = mg.GpuServer(scanner, m, segments=4, lane_depth=4, max_delta_entries=3000,
                       double_buffer=double_buffer)
    rng = np.random.default_rng(7)
    state = rows.copy()
    expected = []
    for epoch in range(1, 6):
        keys = make_keys(d, [int(x) for x in rng.integers(0, n, size=6)], epoch % 2, 50 * epoch)
        expected.append((keys, srv.submit(keys), state.copy(), epoch - 1))
        idx = rng.integers(0, n, size=1000).astype(np.uint64)
        idx[:3] = idx[3]
        diffs = rng.integers(0, 256, size=(1000, w), dtype=np.uint8)
        srv.update(idx, diffs, epoch)
        for i, dlt in zip(idx, diffs):
            state[int(i)] ^= dlt
    keys = make_keys(d, [0, n - 1, 12345], 1, 999)
    expected.append((keys, srv.submit(keys), state.copy(), 5))
    for keys, t, snap, epoch in expected:
        ans, info = srv.wait(t, len(keys))
        assert info["epoch"] == epoch
        assert np.array_equal(ans, oracle.scan_tiled(keys, snap, THREADS)), epoch
    st = srv.stats()
    assert st["epoch"] == 5 and st["rebases"] >= 1
    srv.close()
    # After the loop stops the caller's matrix holds the merged rows (rows as
    # of the last merge; later updates live only in the delta).
    m.close()


def test_server_rejects_bad_requests(mg, scanner):
    n, w = 1 << 12, 32
    m = mg.GpuPageMatrix(scanner, n, w)
    m.fill_seeded(3)
    srv = mg.GpuServer(scanner, m, segments=2, lane_depth=2, max_active_keys=6)
    with pytest.raises(mg.InvalidDomainBits):
        srv.submit([oracle.keygen_bytes(13, 1, seed=1)[0]])
    with pytest.raises(mg.MgError) as e:
        srv.submit([oracle.keygen_bytes(12, 1, seed=i)[0] for i in range(7)])
    assert e.value.code == mg.MG_ERR_BUSY
    with pytest.raises(mg.RowsError):
        srv.update(np.array([n], dtype=np.uint64), np.zeros((1, w), np.uint8), 1)
    keys = [oracle.keygen_bytes(12, 5, seed=2)[0]]
    t = srv.submit(keys)
    ans, _ = srv.wait(t, 1)
    assert np.array_equal(ans, scanner.scan_keys(m, keys))
    with pytest.raises(mg.MgError):
        srv.wait(t, 1)  # already collected
    srv.close()
    m.close()


def test_wide_rows_through_the_server(mg, scanner):
    n, w, d = 5000, 1024, 13
    rows = oracle.fill_rows(64, 0, n, w)
    m = mg.GpuPageMatrix.new(scanner, rows)
    srv = mg.GpuServer(scanner, m, segments=3)
    keys = make_keys(d, [0, 4999, 2500], 0, 5)
    t = srv.submit(keys)
    ans, _ = srv.wait(t, 3)
    assert np.array_equal(ans, oracle.scan_tiled(keys, rows, THREADS))
    srv.close()
    m.close()


def test_lifetimes_are_enforced(mg):
    # A loop's matrix cannot be freed under it, nor the device under its objects.
    scanner = mg.GpuScanner(0)
    m = mg.GpuPageMatrix(scanner, 1 << 12, 32)
    spare = mg.GpuPageMatrix(scanner, 1 << 12, 32)
    srv = mg.GpuServer(scanner, m, segments=2, lane_depth=2, max_active_keys=6)
    with pytest.raises(mg.MgError) as e:
        m.close()
    assert e.value.code == mg.MG_ERR_STATE
    with pytest.raises(mg.MgError) as e:
        scanner.close()
    assert e.value.code == mg.MG_ERR_STATE
    srv.submit_snapshot(spare, 1)      # the loop now serves `spare`: `m` is free
    m.close()
    with pytest.raises(mg.MgError):
        spare.close()
    keys = [oracle.keygen_bytes(12, 7, seed=1)[0]]
    ans, info = srv.wait(srv.submit(keys), 1)
    assert info["epoch"] == 1 and np.array_equal(ans, scanner.scan_keys(spare, keys))
    srv.close()
    spare.close()
    scanner.close()


def test_abandoned_requests_are_dropped(mg, scanner):
    n, d = 1 << 12, 12
    m = mg.GpuPageMatrix(scanner, n, 32)
    m.fill_seeded(4)
    srv = mg.GpuServer(scanner, m, segments=4, lane_depth=2)
    keys = [oracle.keygen_bytes(d, t, seed=t)[0] for t in (1, 2, 3)]
    t_gone = srv.submit(keys)
    srv.abandon(t_gone)                  # still running; its answers will be dropped
    t_kept = srv.submit(keys)
    ans, _ = srv.wait(t_kept, 3)
    assert np.array_equal(ans, scanner.scan_keys(m, keys))
    deadline = time.time() + 30
    while srv.stats()["completed"] < 2 and time.time() < deadline:
        time.sleep(0.01)
    assert srv.stats()["completed"] == 2
    with pytest.raises(mg.MgError):
        srv.wait(t_gone, 3, timeout_ms=100)  # no longer known
    t_done = srv.submit(keys)
    time.sleep(0.2)
    srv.abandon(t_done)                  # done or not, the ticket goes away
    with pytest.raises(mg.MgError):
        srv.abandon(t_done + 1000)       # unknown ticket
    srv.close()
    m.close()
