"""Concurrent calls through one device handle (the reference shares one
GpuScanner across threads over a ticket-based pool of workspace slots,
crates/morphogen-gpu-dpf/src/kernel.rs:50-72, 733-739, and deals per-query
launches over forked streams, :955-1097).

Scans and key evaluations from several threads run on distinct workspaces
(streams); every answer is compared with the oracle. Matrix writes running
beside scans must never leave a scan with a half-written matrix: each answer
equals the oracle on one of the two matrix versions.
"""
import os
import threading

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


def keys_for(d, targets, seed):
    return [oracle.keygen_bytes(d, t, seed=seed + 31 * i)[i % 2] for i, t in enumerate(targets)]


def run_threads(fns):
    errors = []

    def wrap(f):
        try:
            f()
        except BaseException as e:  # reported after the join
            errors.append(e)

    ts = [threading.Thread(target=wrap, args=(f,)) for f in fns]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=300)
    assert not any(t.is_alive() for t in ts), "a thread did not finish"
    if errors:
        raise errors[0]


def test_scans_from_many_threads_match_oracle(mg, scanner):
    n, w, d = 40_000, 64, 16
    print(f"case threads n={n} w={w} d={d}", flush=True)
    rows = oracle.fill_rows(seed=11, first_row=0, n_rows=n, row_bytes=w)
    m = mg.GpuPageMatrix.new(scanner, rows)
    rng = np.random.default_rng(3)
    jobs = []
    for j in range(12):
        q = int(rng.integers(1, 7))
        keys = keys_for(d, [int(x) for x in rng.integers(0, n, size=q)], seed=100 * j)
        jobs.append((keys, oracle.scan_tiled(keys, rows, THREADS)))
    opts = [None, mg.make_opts(kernel=mg.KERNEL_STREAM), mg.make_opts(use_graph=True),
            mg.make_opts(kernel=mg.KERNEL_STREAM_SPLIT, lane_depth=4)]
    for n_ws in (1, 3, 8):
        scanner.set_workspaces(n_ws)

        def worker(t):
            def f():
                for r in range(3):
                    keys, want = jobs[(t + r) % len(jobs)]
                    got = scanner.scan_keys(m, keys, opts[(t + r) % len(opts)])
                    assert np.array_equal(got, want), (n_ws, t, r)
            return f

        run_threads([worker(t) for t in range(8)])
    scanner.set_workspaces(4)
    m.close()


def test_key_evaluation_beside_scans(mg, scanner):
    n, w, d = 8192, 32, 13
    rows = oracle.fill_rows(seed=5, first_row=0, n_rows=n, row_bytes=w)
    m = mg.GpuPageMatrix.new(scanner, rows)
    keys = keys_for(d, [1, 77, 4000, 8191], seed=9)
    want_scan = oracle.scan_tiled(keys, rows, THREADS)
    k0 = oracle.key_from_bytes(keys[0])
    want_full = oracle.full_eval(k0).view(np.uint8).reshape(-1, 16)  # little-endian words
    idx = np.array([0, 5, 8191, 1234], dtype=np.uint64)

    def scans():
        for _ in range(20):
            assert np.array_equal(scanner.scan_keys(m, keys), want_scan)

    def evals():
        for _ in range(10):
            assert np.array_equal(scanner.full_eval(keys[0]), want_full)
            assert np.array_equal(scanner.eval(keys[0], idx), want_full[idx.astype(np.int64)])

    run_threads([scans, scans, evals, evals])
    m.close()


def test_per_query_over_workspaces_equals_one_pass(mg, scanner):
    n, w, d = 30_000, 256, 15
    print(f"case per-query n={n} w={w} d={d}", flush=True)
    rows = oracle.fill_rows(seed=21, first_row=0, n_rows=n, row_bytes=w)
    m = mg.GpuPageMatrix.new(scanner, rows)
    rng = np.random.default_rng(8)
    keys = keys_for(d, [int(x) for x in rng.integers(0, n, size=3 * 11)], seed=40)
    want = oracle.scan_tiled(keys, rows, THREADS)
    assert np.array_equal(scanner.scan_keys(m, keys), want)
    for n_ws in (1, 2, 4, 8):
        scanner.set_workspaces(n_ws)
        for per in (3, 5, 33, 40):
            for graph in (False, True):
                got, t = scanner.scan_keys_per_query(m, keys, per, mg.make_opts(use_graph=graph),
                                                     with_timing=True)
                assert np.array_equal(got, want), (n_ws, per, graph)
                assert t["total_ms"] > 0
    # Two callers dealing their queries over the same workspaces at once.
    scanner.set_workspaces(4)

    def caller():
        for _ in range(4):
            assert np.array_equal(scanner.scan_keys_per_query(m, keys, 3), want)

    run_threads([caller, caller, caller])
    with pytest.raises(mg.MgError):
        scanner.set_workspaces(0)
    with pytest.raises(mg.MgError):
        scanner.set_workspaces(9)
    m.close()


def test_matrix_writes_beside_scans_are_never_seen_half_done(mg, scanner):
    n, w, d = 1 << 16, 128, 16
    print(f"case writes n={n} w={w} d={d}", flush=True)
    va = oracle.fill_rows(seed=1, first_row=0, n_rows=n, row_bytes=w)
    vb = oracle.fill_rows(seed=2, first_row=0, n_rows=n, row_bytes=w)
    m = mg.GpuPageMatrix.new(scanner, va)
    keys = keys_for(d, [0, n - 1, 12345], seed=77)
    want_a = oracle.scan_tiled(keys, va, THREADS)
    want_b = oracle.scan_tiled(keys, vb, THREADS)
    stop = threading.Event()
    seen = {"a": 0, "b": 0}

    def writer():
        try:
            for r in range(30):
                m.update_rows(0, vb if r % 2 == 0 else va)
        finally:
            stop.set()

    def reader():
        while not stop.is_set():
            got = scanner.scan_keys(m, keys)
            if np.array_equal(got, want_a):
                seen["a"] += 1
            else:
                assert np.array_equal(got, want_b), "a scan saw a half-written matrix"
                seen["b"] += 1

    run_threads([writer, reader, reader, reader])
    assert seen["a"] + seen["b"] > 0
    m.close()
