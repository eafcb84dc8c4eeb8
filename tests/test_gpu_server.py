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
    states = [state.copy()]  # the matrix at each epoch
    expected = []
    for epoch in range(1, 6):
        keys = make_keys(d, [int(x) for x in rng.integers(0, n, size=6)], epoch % 2, 50 * epoch)
        expected.append((keys, srv.submit(keys), epoch - 1))
        idx = rng.integers(0, n, size=1000).astype(np.uint64)
        idx[:3] = idx[3]
        diffs = rng.integers(0, 256, size=(1000, w), dtype=np.uint8)
        srv.update(idx, diffs, epoch)
        for i, dlt in zip(idx, diffs):
            state[int(i)] ^= dlt
        states.append(state.copy())
    keys = make_keys(d, [0, n - 1, 12345], 1, 999)
    expected.append((keys, srv.submit(keys), 5))
    # A request answers for the epoch it was admitted at, never an older one than
    # at its submission. Usually that is the submission's epoch; an update behind
    # a request that waits (for slots, or for a single-buffered merge) overtakes
    # it, and the request then answers at the newer epoch.
    for keys, t, submitted in expected:
        ans, info = srv.wait(t, len(keys))
        e = info["epoch"]
        assert submitted <= e <= 5
        assert np.array_equal(ans, oracle.scan_tiled(keys, states[e], THREADS)), (submitted, e)
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
    srv.abandon(t_gone)                  # queued or in flight: it stops costing GPU time
    t_kept = srv.submit(keys)
    ans, _ = srv.wait(t_kept, 3)
    assert np.array_equal(ans, scanner.scan_keys(m, keys))
    st = srv.stats()
    assert st["abandoned"] == 1 and 1 <= st["completed"] <= 2
    with pytest.raises(mg.MgError):
        srv.wait(t_gone, 3, timeout_ms=100)  # no longer known
    t_done = srv.submit(keys)
    time.sleep(0.2)
    srv.abandon(t_done)                  # done or not, the ticket goes away
    with pytest.raises(mg.MgError):
        srv.abandon(t_done + 1000)       # unknown ticket
    srv.close()
    m.close()


def wait_until(pred, seconds=60):
    end = time.time() + seconds
    while not pred():
        assert time.time() < end, "condition not reached"
        time.sleep(0.002)


def test_abandon_removes_queued_and_frees_in_flight_slots(mg, scanner):
    # 64 key slots: a 48-key request runs (a pass over 2^28 rows takes a good
    # fraction of a second), a second 48-key request must queue behind it.
    n, w, d, seed = 1 << 28, 32, 28, 71
    m = mg.GpuPageMatrix(scanner, n, w)
    m.fill_seeded(seed)
    srv = mg.GpuServer(scanner, m, segments=64, max_active_keys=64)
    rng = np.random.default_rng(71)
    big = make_keys(d, [int(x) for x in rng.integers(0, n, size=48)], 0, 700)
    t_run = srv.submit(big)
    wait_until(lambda: srv.stats()["active_keys"] == 48)
    t_queued = srv.submit(make_keys(d, [int(x) for x in rng.integers(0, n, size=48)], 1, 800))
    assert srv.stats()["queued_requests"] == 1
    srv.abandon(t_queued)                # never admitted: it leaves the queue
    st = srv.stats()
    assert st["queued_requests"] == 0 and st["abandoned"] == 1
    srv.abandon(t_run)                   # in flight: its keys leave at the next step
    wait_until(lambda: srv.stats()["active_keys"] == 0)
    st = srv.stats()
    assert st["completed"] == 0 and st["abandoned"] == 2 and st["keys_completed"] == 0
    assert st["sweep_steps"] < 64        # it left before finishing its pass
    # The slots are free again and the loop answers correctly.
    keys = make_keys(d, [n // 3, 17, n - 2], 1, 900)
    ans, _ = srv.wait(srv.submit(keys), 3)
    assert np.array_equal(ans, oracle.scan_seeded(keys, seed, n, w, n_threads=THREADS))
    srv.close()
    m.close()


def test_admission_deadline_refuses_late_requests(mg, scanner):
    n, w, d, seed = 1 << 24, 32, 24, 72
    m = mg.GpuPageMatrix(scanner, n, w)
    m.fill_seeded(seed)

    def measured_key_step(srv):
        # Before any step has been timed every request is admitted; one request
        # gives the loop its time per key and sweep step.
        warm = make_keys(d, [1, 2, 3], 0, 1)
        srv.wait(srv.submit(warm), 3)
        return srv.stats()["key_step_ms"]

    probe = mg.GpuServer(scanner, m, segments=16)
    per_key_step = measured_key_step(probe)
    probe.close()
    assert per_key_step > 0
    # A deadline worth about 60 keys of backlog, then a burst of 720 keys: the
    # late part is refused at submit with MG_ERR_BUSY (the wire layer's 503),
    # the admitted part is answered exactly.
    deadline = max(2, int(per_key_step * 16 * 60))
    srv = mg.GpuServer(scanner, m, segments=16, deadline_ms=deadline)
    measured_key_step(srv)
    rng = np.random.default_rng(3)
    accepted, refused = [], 0
    for r in range(60):
        keys = make_keys(d, [int(x) for x in rng.integers(0, n, size=12)], r % 2, 1000 + 20 * r)
        try:
            accepted.append((keys, srv.submit(keys)))
        except mg.MgError as e:
            assert e.code == mg.MG_ERR_BUSY
            refused += 1
    st = srv.stats()
    assert refused > 0 and st["rejected"] == refused and accepted
    lat = []
    for keys, t in accepted:
        ans, info = srv.wait(t, len(keys))
        lat.append(info["latency_ms"])
        assert np.array_equal(ans, scanner.scan_keys(m, keys))
    ref = oracle.scan_seeded(accepted[0][0][:2], seed, n, w, n_threads=THREADS)
    assert np.array_equal(srv_answers_first := scanner.scan_keys(m, accepted[0][0][:2]), ref)
    print(f"deadline {deadline} ms: {len(accepted)} accepted, {refused} refused, "
          f"max latency {max(lat):.0f} ms, key-step {per_key_step:.4f} ms", flush=True)
    srv.close()
    m.close()


def test_update_applies_ahead_of_queued_requests(mg, scanner):
    # The update is submitted after a request that waits for slots; it lands at
    # the next sweep step, so that request answers at the newer epoch (and says
    # so), while the request in flight keeps the epoch it was admitted at.
    n, w, d, seed = 1 << 28, 32, 28, 73
    m = mg.GpuPageMatrix(scanner, n, w)
    m.fill_seeded(seed)
    srv = mg.GpuServer(scanner, m, segments=64, max_active_keys=48)
    rng = np.random.default_rng(73)
    first = make_keys(d, [int(x) for x in rng.integers(0, n, size=48)], 0, 10)
    t1 = srv.submit(first)
    wait_until(lambda: srv.stats()["active_keys"] == 48)
    second = make_keys(d, [10, 11, 12], 1, 20)
    t2 = srv.submit(second)
    idx = np.array([10, 11, 500], dtype=np.uint64)
    diffs = rng.integers(0, 256, size=(3, w), dtype=np.uint8)
    srv.update(idx, diffs, 1)
    wait_until(lambda: srv.stats()["epoch"] == 1)
    assert srv.stats()["completed"] == 0      # the update did not wait for the first request
    a1, i1 = srv.wait(t1, 48)
    a2, i2 = srv.wait(t2, 3)
    assert i1["epoch"] == 0 and i2["epoch"] == 1
    j = [0, 47]
    assert np.array_equal(a1[j], oracle.scan_seeded([first[i] for i in j], seed, n, w,
                                                    n_threads=THREADS))
    want2 = (oracle.scan_seeded(second, seed, n, w, n_threads=THREADS)
             ^ oracle.scan_delta(second, idx, diffs))
    assert np.array_equal(a2, want2)
    srv.close()
    m.close()


@pytest.mark.parametrize("w,segments", [(32, 8), (4096, 2)])
def test_graph_replayed_sweep_steps_equal_oracle(mg, scanner, w, segments):
    n, d = (1 << 17, 17) if w == 32 else (9000, 14)
    rows = oracle.fill_rows(74 + w, 0, n, w)
    m = mg.GpuPageMatrix.new(scanner, rows)
    srv = mg.GpuServer(scanner, m, segments=segments, use_graph=True)
    rng = np.random.default_rng(w)
    reqs = [make_keys(d, [int(x) for x in rng.integers(0, n, size=k)], r % 2, 100 * r)
            for r, k in enumerate([3, 6, 4, 3, 9])]
    # Twice the same requests: the second round replays the captured graphs.
    for rnd in range(2):
        tickets = [srv.submit(keys) for keys in reqs]
        for keys, t in zip(reqs, tickets):
            ans, _ = srv.wait(t, len(keys))
            assert np.array_equal(ans, oracle.scan_tiled(keys, rows, THREADS)), rnd
    assert srv.stats()["graph_launches"] > 0
    srv.close()
    m.close()


def test_wide_pages_in_column_halves_through_the_server(mg, scanner):
    # 4 KiB rows from 4 keys per step run as two column halves per row; the
    # loop's segments, slot lists and accumulation must see the same answers.
    n, w, d = 9000, 4096, 14
    rows = oracle.fill_rows(75, 0, n, w)
    m = mg.GpuPageMatrix.new(scanner, rows)
    srv = mg.GpuServer(scanner, m, segments=3)
    keys = make_keys(d, [0, 8999, 4500, 123, 7777], 1, 15)
    ans, _ = srv.wait(srv.submit(keys), 5)
    assert np.array_equal(ans, oracle.scan_tiled(keys, rows, THREADS))
    # Requests of 9, 4 and 3 keys submitted together: steps of up to 16 keys run as several
    # groups of 4, paced, with both column halves of a group folding one DFS wave's masks.
    rng = np.random.default_rng(8)
    reqs = [make_keys(d, [int(x) for x in rng.integers(0, n, size=k)], k % 2, 40 + 7 * k)
            for k in (9, 4, 3)]
    tickets = [srv.submit(r) for r in reqs]
    for t, r in zip(tickets, reqs):
        ans, _ = srv.wait(t, len(r))
        assert np.array_equal(ans, oracle.scan_tiled(r, rows, THREADS))
    srv.close()
    m.close()


def test_injected_failure_reaches_waiters_and_queued_snapshot(mg, scanner):
    # The failure path: a request in flight and a snapshot queued behind it both
    # get the error; later calls are refused; destroying the loop is safe.
    n, w, d = 1 << 28, 32, 28
    m = mg.GpuPageMatrix(scanner, n, w)
    m.fill_seeded(5)
    spare = mg.GpuPageMatrix(scanner, n, w)
    spare.fill_seeded(6)
    srv = mg.GpuServer(scanner, m, segments=64)
    rng = np.random.default_rng(76)
    keys = make_keys(d, [int(x) for x in rng.integers(0, n, size=48)], 0, 3)
    t = srv.submit(keys)
    wait_until(lambda: srv.stats()["active_keys"] == 48)
    snap_err = []

    def swap():
        try:
            srv.submit_snapshot(spare, 9)
        except mg.MgError as e:
            snap_err.append(e)

    th = threading.Thread(target=swap)
    th.start()
    time.sleep(0.05)                      # the snapshot queues behind the keys in flight
    assert srv.stats()["completed"] == 0
    srv.inject_failure(mg.MG_ERR_DEVICE)
    with pytest.raises(mg.MgError) as e:
        srv.wait(t, 48)
    assert e.value.code == mg.MG_ERR_DEVICE
    th.join(timeout=60)
    assert not th.is_alive() and len(snap_err) == 1
    with pytest.raises(mg.MgError):
        srv.submit(make_keys(d, [4], 0, 4))
    srv.close()
    m.close()
    spare.close()
