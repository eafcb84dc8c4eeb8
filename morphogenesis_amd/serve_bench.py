"""Serving benchmark: Poisson request arrivals against the request loop.

Each request carries 3 cuckoo keys x k lookups, k uniform in [k_min, k_max];
requests arrive as a Poisson process at --rate per second for --duration
seconds. An epoch update of --delta-rows uniform rows (random new values)
is applied every --epoch-seconds. Keys come in party-0/party-1 pairs, so every
completed request is checked bit for bit: each pair's answers must XOR to
the target row as of the request's admission epoch (seeded base row XOR the
updates up to that epoch, recomputed on the host).

Reported (one JSON line): requests admitted, refused at submit by the
admission deadline (--deadline-ms) and completed, queries/s (3 keys = 1
query), p50/p99 latency (submit -> answers on the host) of completed
requests and the share within the deadline, requests still in flight at the
end, how long each epoch update took to reach the loop, and the plain rate
(one 3-key query per full pass) measured on the same matrix.

CLI: python -m morphogenesis_amd.serve_bench --rows-log2 27 --rate 5 --duration 20
"""
import argparse
import json
import sys
import threading
import time

import numpy as np

from . import MG_ERR_BUSY, GpuPageMatrix, GpuScanner, GpuServer, MgError, generate_keys
from .data import seeded_rows


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(xs, p):
    return float(np.percentile(np.asarray(xs), p)) if xs else None


def run(scanner, m, seed, rate=5.0, duration=20.0, drain=30.0, k_min=1, k_max=32,
        epoch_seconds=12.0, delta_rows=1 << 16, segments=0, max_active_keys=96,
        double_buffer=False, deadline_ms=0, use_graph=False, rng_seed=None, log=log):
    """Poisson arrivals against a request loop over `m` (rows seeded with `seed`),
    every answer checked; returns the report dict (see the module docstring)."""
    n, w = m.n_rows, m.row_bytes
    d = max(1, (n - 1).bit_length())
    rng = np.random.default_rng(seed if rng_seed is None else rng_seed)

    # Plain path: one query (3 keys) per full pass over the same matrix.
    plain_keys = [generate_keys(d, int(t), seed=90 + i)[0]
                  for i, t in enumerate(rng.integers(0, n, size=3))]
    scanner.scan_keys(m, plain_keys)
    t0 = time.perf_counter()
    scanner.scan_keys(m, plain_keys)
    plain_ms = (time.perf_counter() - t0) * 1e3
    log(f"plain pass (3 keys): {plain_ms:.1f} ms")

    srv = GpuServer(scanner, m, segments=segments, max_active_keys=max_active_keys,
                    max_delta_entries=max(1 << 20, 8 * delta_rows),
                    double_buffer=double_buffer, deadline_ms=deadline_ms, use_graph=use_graph)
    log(f"server: {srv.stats()}")

    lock = threading.Lock()
    updates = {}   # row -> [(epoch, diff)] in epoch order
    records = []   # (ticket, targets, n_keys)
    refused = []   # n_keys of requests refused at submit (admission deadline)
    update_lag_ms = []  # update call -> epoch visible in the loop's stats
    stop = threading.Event()

    rng_s = np.random.default_rng(rng.integers(1 << 62))  # one generator per thread
    rng_e = np.random.default_rng(rng.integers(1 << 62))

    def submitter():
        r = rng_s
        t_next = time.perf_counter()
        i = 0
        while not stop.is_set():
            t_next += r.exponential(1.0 / rate)
            delay = t_next - time.perf_counter()
            if delay > 0 and stop.wait(delay):
                break
            n_keys = 3 * int(r.integers(k_min, k_max + 1))
            targets = [int(x) for x in r.integers(0, n, size=(n_keys + 1) // 2)]
            keys = []
            for j, t in enumerate(targets):
                keys += list(generate_keys(d, t, seed=(i << 16) + j))
            keys = keys[:n_keys]
            with lock:
                try:
                    records.append((srv.submit(keys), targets, n_keys))
                except MgError as e:
                    if e.code != MG_ERR_BUSY:
                        raise
                    refused.append(n_keys)
            i += 1

    def epocher():
        r = rng_e
        e = 0
        while not stop.wait(epoch_seconds):
            e += 1
            idx = r.integers(0, n, size=delta_rows).astype(np.uint64)
            diffs = r.integers(0, 256, size=(delta_rows, w), dtype=np.uint8)
            with lock:
                t_up = time.perf_counter()
                srv.update(idx, diffs, e)
                for row, dlt in zip(idx.tolist(), diffs):
                    updates.setdefault(row, []).append((e, dlt.copy()))
            while srv.stats()["epoch"] < e and not stop.is_set():
                time.sleep(0.001)
            if srv.stats()["epoch"] >= e:
                update_lag_ms.append((time.perf_counter() - t_up) * 1e3)
                log(f"epoch {e}: {delta_rows} rows updated, in the loop after "
                    f"{update_lag_ms[-1]:.1f} ms")

    th_s = threading.Thread(target=submitter)
    th_e = threading.Thread(target=epocher)
    t_start = time.perf_counter()
    th_s.start()
    th_e.start()
    time.sleep(duration)
    stop.set()
    th_s.join()
    th_e.join()
    log(f"arrivals done: {len(records)} requests admitted, {len(refused)} refused")

    lat, queue_ms = [], []
    done = keys_done = unfinished = 0
    exact = True
    deadline = time.perf_counter() + drain
    t_last = time.perf_counter()
    for ticket, targets, n_keys in records:
        remaining = int((deadline - time.perf_counter()) * 1e3)
        if remaining <= 0:
            unfinished += 1
            continue
        try:
            ans, info = srv.wait(ticket, n_keys, timeout_ms=remaining)
        except MgError:
            unfinished += 1
            continue
        t_last = time.perf_counter()
        done += 1
        keys_done += n_keys
        lat.append(info["latency_ms"])
        queue_ms.append(info["queue_ms"])
        for j, t in enumerate(targets):
            if 2 * j + 1 >= n_keys:
                break
            want = seeded_rows(seed, t, 1, w)[0].copy()
            for e, dlt in updates.get(t, []):
                if e <= info["epoch"]:
                    want ^= dlt
            if not np.array_equal(ans[2 * j] ^ ans[2 * j + 1], want):
                exact = False
    wall = t_last - t_start
    st = srv.stats()
    n_rows_label = f"2^{n.bit_length() - 1}" if n & (n - 1) == 0 else str(n)
    out = {
        "workload": f"{n_rows_label} rows x {w} B, Poisson {rate} req/s for {duration} s, "
                    f"3 x U{{{k_min}..{k_max}}} keys per request, {delta_rows}-row epoch update "
                    f"every {epoch_seconds} s"
                    + (f", admission deadline {deadline_ms} ms" if deadline_ms else ""),
        "submitted": len(records) + len(refused), "admitted": len(records),
        "refused": len(refused),
        "refused_frac": len(refused) / max(1, len(records) + len(refused)),
        "completed": done, "unfinished": unfinished,
        "keys_completed": keys_done,
        "queries_per_s": keys_done / 3.0 / wall if wall > 0 else None,
        "keys_per_s": keys_done / wall if wall > 0 else None,
        "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "max_ms": max(lat) if lat else None,
        "p50_queue_ms": pct(queue_ms, 50),
        "deadline_ms": deadline_ms or None,
        "met_deadline_frac": (sum(1 for x in lat if x <= deadline_ms) / len(lat))
        if (deadline_ms and lat) else None,
        "offered_keys_per_s": rate * 3 * (k_min + k_max) / 2.0,
        "plain_queries_per_s": 1e3 / plain_ms,
        "update_lag_ms": update_lag_ms,
        "sweep_step_ms": st["busy_ms"] / st["sweep_steps"] if st["sweep_steps"] else None,
        "use_graph": bool(use_graph),
        "bit_exact": exact,
        "server": st,
    }
    srv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rows-log2", type=int, default=27)
    ap.add_argument("--row-bytes", type=int, default=32)
    ap.add_argument("--rate", type=float, default=5.0, help="requests per second")
    ap.add_argument("--duration", type=float, default=20.0, help="seconds of arrivals")
    ap.add_argument("--drain", type=float, default=30.0, help="seconds to wait for stragglers")
    ap.add_argument("--k-min", type=int, default=1)
    ap.add_argument("--k-max", type=int, default=32)
    ap.add_argument("--epoch-seconds", type=float, default=12.0)
    ap.add_argument("--delta-rows", type=int, default=1 << 16)
    ap.add_argument("--segments", type=int, default=0)
    # Admission is bounded by free HBM; this cap additionally bounds the keys
    # sharing a sweep, trading throughput headroom for latency (FIFO beyond it).
    ap.add_argument("--max-active-keys", type=int, default=96)
    ap.add_argument("--double-buffer", action="store_true")
    ap.add_argument("--deadline-ms", type=int, default=0,
                    help="admission deadline: refuse requests predicted to finish later")
    ap.add_argument("--use-graph", action="store_true",
                    help="replay each sweep step from a HIP graph")
    ap.add_argument("--seed", type=int, default=5)
    args = ap.parse_args()

    if args.max_active_keys < 3 * args.k_max:
        raise SystemExit(f"--max-active-keys {args.max_active_keys} is below the largest request "
                         f"({3 * args.k_max} keys); such requests would be rejected")
    n = 1 << args.rows_log2
    scanner = GpuScanner(0)
    m = GpuPageMatrix(scanner, n, args.row_bytes)
    m.fill_seeded(args.seed)
    out = run(scanner, m, args.seed, rate=args.rate, duration=args.duration, drain=args.drain,
              k_min=args.k_min, k_max=args.k_max, epoch_seconds=args.epoch_seconds,
              delta_rows=args.delta_rows, segments=args.segments,
              max_active_keys=args.max_active_keys, double_buffer=args.double_buffer,
              deadline_ms=args.deadline_ms, use_graph=args.use_graph)
    print(json.dumps(out))
    m.close()
    scanner.close()


if __name__ == "__main__":
    main()
