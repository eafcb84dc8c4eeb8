"""Benchmark of the fused ChaCha8-DPF masked-XOR PIR scan on one MI355X.

Headline workload: the full-mainnet shape, 2^31 rows x 32 B (64 GiB) of
seeded uniform bytes resident in HBM plus an epoch DeltaBuffer of 2^16
updated rows (uniform indices, random diffs), DPF keys over a 31-bit domain
with uniform targets. One step = one batched pass of --keys keys over the
whole matrix and the delta. Keys come in party-0/party-1 pairs so every step
also checks, bit for bit, that each pair's answers XOR to the target row of
the updated matrix.

Reported on the same JSON line:
  - batched queries/s (a query = 3 cuckoo keys), per-pass p50/p99, the plain
    one-query-per-pass rate and the reference-shaped plain kernel beside it,
    scan GB/s, and the roofline of the fused kernel (HIP-event time on the
    kernel's stream);
  - by_keys: single passes at 3, 96 and 384 keys on the same matrix;
  - graph: the headline pass replayed from a HIP graph against eager launches;
  - by_config: the other workloads, each with its roofline fraction and a
    bit-exact flag: 2^20 x 32 B (the CPU-runnable pair), 2^27 x 32 B (one
    key per server, and the pair in one pass), 78,643,200 x 1 KiB (1 and 3
    keys), and a bounded serving run on the 2^31 x 32 B matrix (Poisson
    arrivals, an admission deadline, an epoch update every 12 s); the
    reference's own 4 KiB page shapes (16,817,159 and 65,536 pages, 3 keys
    per query): one query per pass, 32 queries over the scan workspaces'
    streams, and the same 32 queries batched 2-32 per pass;
  - cpu_baseline: the CPU scan on the host cores of the same run, on a bounded
    sample of the headline workload, in three forms (the reference's tiled
    eval_fused_3dpf_cpu shape, its server scan's eight-rows-per-step shape,
    and a 16-lane vector ChaCha form), with the host's core count, CPU quota,
    model and memory.
"""
import argparse
import json
import os
import platform
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "batched queries/s + p50/p99 ms at 2^31x32B DB; scan GB/s vs HBM peak"
KEYS_PER_QUERY = 3
# MI355X peaks: HBM3E 8 TB/s (published). int32 VALU: 256 CUs x 128 lanes per
# clock x 2.4 GHz = 78.6 TOP/s, the rate of the published 157.3 TFLOP/s FP32
# vector peak counted per op; simple int32 ops (v_add_u32, v_xor_b32) reach it:
# the microbenchmark's independent add/xor chains run at 0.74 of it at the
# 2.1 GHz clock they hold (python -m morphogenesis_amd.microbench --forms).
# Rotates (v_alignbit_b32), a third of ChaCha's ops, issue at 64 lanes per
# clock, so the ChaCha mix (2 of 3 ops at 128, 1 at 64 lanes per clock) is
# bounded at 96 lanes per clock: 59.0 TOP/s, reported beside the peak.
HBM_PEAK_GBS = 8000.0
VALU_PEAK_TOPS = 256 * 128 * 2.4e9 / 1e12
VALU_MIX_PEAK_TOPS = 256 * 96 * 2.4e9 / 1e12
# Algorithmic int32 ops per leaf: one GGM node (2 ChaCha8 blocks after dead-code
# elimination, correction word included) = 700, plus one v_bitop3 per 32-bit
# word of the row for the masked fold.
OPS_PER_NODE = 700
# CPU threads for the oracle legs: the CPU share a one-GPU job gets on the
# GPU hosts (16 of the host's cores, also its cgroup CPU quota there).
CPU_SHARE = 16


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def percentile(xs, p):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * p / 100.0
    lo, hi = int(k), min(int(k) + 1, len(xs) - 1)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def valu_frac(n_rows, w, n_keys, kernel_ms):
    return n_keys * n_rows * (OPS_PER_NODE + w // 4) / (kernel_ms * 1e-3) / 1e12 / VALU_PEAK_TOPS


def hbm_frac(n_rows, w, kernel_ms):
    return n_rows * w / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS


def cpu_threads():
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, min(n, CPU_SHARE))


def host_info():
    info = {"os_cpu_count": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "machine": platform.machine()}
    try:
        with open("/proc/cpuinfo") as f:
            info["cpu_model"] = next(l.split(":", 1)[1].strip() for l in f
                                     if l.startswith("model name"))
    except Exception:
        pass
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()
            info["cgroup_cpu_quota"] = None if q == "max" else int(q) / int(period)
    except Exception:
        info["cgroup_cpu_quota"] = None
    try:
        with open("/proc/meminfo") as f:
            kb = next(int(l.split()[1]) for l in f if l.startswith("MemTotal"))
        info["mem_total_gib"] = round(kb / 2**20, 1)
    except Exception:
        pass
    return info


def cpu_baseline(n_rows, row_bytes, domain_bits, sample_rows_log2, seed):
    """The CPU scan of one query (3 keys) on a prefix of the same matrix, on the
    host cores of this run, in three forms; throughput scaled to the full row
    count. value = the fastest form."""
    import numpy as np

    import oracle
    threads = cpu_threads()
    n = 1 << sample_rows_log2
    rows = oracle.fill_rows(seed, 0, n, row_bytes)
    keys = [oracle.keygen_bytes(domain_bits, (i * 2654435761) % n, seed=100 + i)[0]
            for i in range(KEYS_PER_QUERY)]
    forms = [
        ("tiled", "eval_fused_3dpf_cpu's shape (kernel.rs:1276-1371): 2048-row tiles handed "
                  "to threads, scalar ChaCha8 eval_subtree per key and tile, masked XOR",
         lambda: oracle.scan_tiled(keys, rows, threads)),
        ("rows8", "the server scan's shape (scan.rs:600-649 fold over chunks, :721-895 eight "
                  "rows x 64 B per step in 512-bit vectors) with scalar ChaCha8 masks",
         lambda: oracle.scan_rows8(keys, rows, threads)),
        ("simd", "16 tree nodes per ChaCha8 evaluation in 512-bit vectors, breadth-first "
                 "4096-leaf subtrees, 512-bit masked folds (oracle_fast.c)",
         lambda: oracle.scan_fast(keys, rows, threads)),
    ]
    ref = None
    variants = []
    full_scale = n_rows / n
    for name, what, fn in forms:
        fn()  # warm: page in the rows and the code
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            ans = fn()
            times.append(time.perf_counter() - t0)
        if ref is None:
            ref = ans
        elif not np.array_equal(ans, ref):
            raise SystemExit(f"cpu baseline form {name} differs from the tiled form")
        t = statistics.median(times)
        variants.append({"form": name, "what": what, "seconds_per_sample": t,
                         "queries_per_s": 1.0 / (t * full_scale)})
        log(f"cpu {name}: {t:.2f} s per sample ({1.0 / (t * full_scale):.4f} queries/s)")
    best = max(variants, key=lambda v: v["queries_per_s"])
    host = host_info()
    return {"value": best["queries_per_s"], "unit": "queries/s", "cores": threads,
            "kind": "port",
            "sample": f"3 keys (1 query) over the first 2^{sample_rows_log2} of {n_rows} rows x "
                      f"{row_bytes} B, median of 3 passes, scaled x{full_scale:.0f} to the full "
                      f"matrix; value = fastest form ({best['form']}); {threads} threads "
                      f"(the job's CPU share) on {host.get('cpu_model', 'unknown CPU')}",
            "isa": oracle.simd_isa(), "variants": variants, "host": host}


def scan_passes(scanner, m, keys, opts, reps, warm=1):
    ans = None
    kern, wall = [], []
    t = None
    for i in range(warm + reps):
        s0 = time.perf_counter()
        ans, t = scanner.scan_keys(m, keys, opts, with_timing=True)
        if i >= warm:
            wall.append((time.perf_counter() - s0) * 1e3)
            kern.append(t["scan_kernel_ms"])
    return ans, kern, wall, t


def config_small(mg, scanner, n_rows, w, d, seed, label, one_key_passes=True):
    """A matrix of n_rows seeded rows and one key pair with a uniform target:
    each server's key alone (and the pair in one pass), every answer compared
    with the oracle (the CPU scan) and the pair's XOR with the target row."""
    import numpy as np

    import oracle
    threads = cpu_threads()
    m = mg.GpuPageMatrix(scanner, n_rows, w)
    m.fill_seeded(seed)
    rng = np.random.default_rng(seed)
    target = int(rng.integers(0, n_rows))
    pair = list(mg.generate_keys(d, target, seed=seed + 7))
    row = oracle.fill_rows(seed, target, 1, w)[0]
    exact = True
    out = {"workload": label, "rows": n_rows, "row_bytes": w, "domain_bits": d}
    per_server = []
    for party in (0, 1):
        # Back-to-back warm passes: after an idle gap the first few short passes run at a
        # lower clock (2^27 x 32 B pair: 5.6 ms after a 50 ms pause, 5.2 ms on the first of a
        # run of passes, 4.6 ms from the tenth on).
        ans, kern, wall, t = scan_passes(scanner, m, [pair[party]], None, 5, warm=10)
        t0 = time.perf_counter()
        ref = oracle.scan_seeded([pair[party]], seed, n_rows, w, n_threads=threads)
        cpu_s = time.perf_counter() - t0
        exact &= bool(np.array_equal(ans, ref))
        per_server.append(ans[0])
        km = statistics.mean(kern)
        out.setdefault("one_key", []).append(
            {"party": party, "kernel_ms": km, "pass_ms": statistics.mean(wall),
             "kernel": t["kernel"], "lane_depth": t["lane_depth"], "waves": t["waves"],
             "gbs": n_rows * w / (km * 1e-3) / 1e9, "hbm_frac": hbm_frac(n_rows, w, km),
             "valu_frac": valu_frac(n_rows, w, 1, km),
             "oracle_s": cpu_s})
    exact &= bool(np.array_equal(per_server[0] ^ per_server[1], row))
    ans, kern, wall, t = scan_passes(scanner, m, pair, None, 5, warm=10)
    exact &= bool(np.array_equal(ans[0] ^ ans[1], row))
    km = statistics.mean(kern)
    out["pair_one_pass"] = {"kernel_ms": km, "pass_ms": statistics.mean(wall),
                            "kernel": t["kernel"], "lane_depth": t["lane_depth"],
                            "waves": t["waves"], "gbs": n_rows * w / (km * 1e-3) / 1e9,
                            "hbm_frac": hbm_frac(n_rows, w, km),
                            "valu_frac": valu_frac(n_rows, w, 2, km),
                            "valu_mix_frac": valu_frac(n_rows, w, 2, km) * VALU_PEAK_TOPS
                            / VALU_MIX_PEAK_TOPS}
    out["bit_exact"] = exact
    m.close()
    return out


def config_pages(mg, scanner, seed, key_counts=(1, 3, 4), placements=2):
    """78,643,200 rows x 1 KiB, 27-bit keys: one pass of k keys per server, each
    server's answers compared with the oracle, pairs XOR-ed to their target rows.

    The matrix is allocated `placements` times side by side (two 75 GiB copies of
    the same rows) and every pass runs on each: where HBM places a matrix moves
    this shape's read rate by up to about 12 % (78,643,200 x 1 KiB, one key: 11.10
    against 11.84 ms for two copies allocated together), so each placement's
    figures are reported and the headline ones are their mean."""
    import numpy as np

    import oracle
    n, w, d = 78_643_200, 1024, 27
    threads = cpu_threads()
    ms = []
    for _ in range(placements):
        m = mg.GpuPageMatrix(scanner, n, w)
        m.fill_seeded(seed)
        ms.append(m)
    rng = np.random.default_rng(seed)
    out = {"workload": "78,643,200 rows x 1 KiB, 27-bit keys, k keys per pass (k cuckoo "
                       "lookups of one query for k = 3)", "rows": n, "row_bytes": w,
           "domain_bits": d, "placements": placements, "by_keys": {}}
    # Untimed passes: the first passes over freshly written rows run slower.
    warm_keys = [mg.generate_keys(d, 1, seed=1)[0]]
    for m in ms:
        for _ in range(2):
            scanner.scan_keys(m, warm_keys)
    for k in key_counts:
        targets = [int(x) for x in rng.integers(0, n, size=k)]
        pairs = [mg.generate_keys(d, t, seed=seed * 10 + 100 * k + i) for i, t in enumerate(targets)]
        exact = True
        per_placement = []
        t = None
        first_answers = None
        for pi, m in enumerate(ms):
            answers = []
            kern_m = []
            for party in (0, 1):
                keys = [p[party] for p in pairs]
                ans, kern, wall, t = scan_passes(scanner, m, keys, None, 5, warm=3)
                kern_m += kern
                if pi == 0:  # the placements hold the same rows: one oracle check per key set
                    exact &= bool(np.array_equal(
                        ans, oracle.scan_seeded(keys, seed, n, w, n_threads=threads)))
                answers.append(ans)
            for q, tg in enumerate(targets):
                exact &= bool(np.array_equal(answers[0][q] ^ answers[1][q],
                                             oracle.fill_rows(seed, tg, 1, w)[0]))
            if pi == 0:
                first_answers = answers[0]
            else:  # every placement answers alike
                exact &= bool(np.array_equal(answers[0], first_answers))
            km = statistics.mean(kern_m)
            per_placement.append({"kernel_ms": km, "kernel_ms_min": min(kern_m),
                                  "hbm_frac": hbm_frac(n, w, km)})
        km = statistics.mean(p["kernel_ms"] for p in per_placement)
        out["by_keys"][k] = {"kernel_ms": km,
                             "gbs": n * w / (km * 1e-3) / 1e9, "hbm_frac": hbm_frac(n, w, km),
                             "by_placement": per_placement,
                             "valu_frac": valu_frac(n, w, k, km), "kernel": t["kernel"],
                             "lane_depth": t["lane_depth"], "keys_per_group": t["keys_per_wave"],
                             "waves": t["waves"], "bit_exact": exact}
        log(f"pages {k} keys: {km:.2f} ms, HBM {hbm_frac(n, w, km):.3f} "
            f"({', '.join('%.3f' % p['hbm_frac'] for p in per_placement)}), exact {exact}")
    for m in ms:
        m.close()
    return out


# The reference's own published shapes (4 KiB pages, 3 keys per query) and its H100 figures
# for them: docs/benchmarks/2026-02-20-task-30-page-gpu-batch-h100-matrix.csv:4 (single
# query on the 16,817,159-page mainnet matrix), :21 (32 queries on 4 streams), and
# docs/benchmarks/2026-02-19-task-24-page-gpu-batch-h100.csv:5 (65,536 pages).
REFERENCE_PAGE_SHAPES = [
    ("pages_16817159x4096", 16_817_159, 25, {"single_query_ms": 30.24,
                                             "four_streams_32_queries_ms": 948.0}),
    ("pages_65536x4096", 65_536, 16, {"single_query_ms": 0.50}),
]


def config_reference_pages(mg, scanner, n, d, ref_h100, seed, n_queries=32):
    """A matrix of n seeded 4 KiB pages, queries of 3 keys (the reference's page
    shape): one query per pass, and n_queries queries dealt over the device's scan
    workspaces (one pass per query, each on its own stream: the reference's
    multistream mode). Bit-exact by two-party recovery of every target page."""
    import numpy as np

    import oracle
    w = 4096
    m = mg.GpuPageMatrix(scanner, n, w)
    m.fill_seeded(seed)
    rng = np.random.default_rng(seed)
    targets = [int(x) for x in rng.integers(0, n, size=3 * n_queries)]
    pairs = [mg.generate_keys(d, t, seed=seed + 13 * i) for i, t in enumerate(targets)]
    keys = [[p[party] for p in pairs] for party in (0, 1)]
    for _ in range(2):  # the first passes over freshly written rows run slower
        scanner.scan_keys(m, keys[0][:3])
    exact = True
    ans1, kern, wall, t = scan_passes(scanner, m, keys[0][:3], None, 10, warm=1)
    ans1b = scanner.scan_keys(m, keys[1][:3])
    for q in range(3):
        exact &= bool(np.array_equal(ans1[q] ^ ans1b[q], oracle.fill_rows(seed, targets[q], 1, w)[0]))
    km, pm = statistics.median(kern), statistics.median(wall)
    out = {"workload": f"{n:,} pages x 4 KiB, {d}-bit keys, 3 keys per query", "rows": n,
           "row_bytes": w, "domain_bits": d,
           "single_query": {"kernel_ms": km, "pass_ms": pm, "queries_per_s": 1e3 / pm,
                            "gbs": n * w / (km * 1e-3) / 1e9, "hbm_frac": hbm_frac(n, w, km),
                            "kernel": t["kernel"], "lane_depth": t["lane_depth"],
                            "waves": t["waves"]},
           "reference_h100": ref_h100}
    walls = []
    answers = None
    for rep in range(3):
        s0 = time.perf_counter()
        a0 = scanner.scan_keys_per_query(m, keys[0], 3)
        walls.append((time.perf_counter() - s0) * 1e3)
        if rep == 0:
            answers = a0
    a1 = scanner.scan_keys_per_query(m, keys[1], 3)
    for q, tg in enumerate(targets):
        exact &= bool(np.array_equal(answers[q] ^ a1[q], oracle.fill_rows(seed, tg, 1, w)[0]))
    bm = statistics.median(walls)
    out["per_query_streams"] = {"queries": n_queries, "workspaces": 4, "batch_ms": bm,
                                "queries_per_s": n_queries * 1e3 / bm,
                                "vs_single_query": (n_queries * 1e3 / bm) / (1e3 / pm)}
    # The same queries batched: qpp queries (3 * qpp keys) per pass, one pass after another,
    # every pass's pages recovered from both parties' answers.
    batched = []
    for qpp in (2, 4, 8, 16, 32):
        per = 3 * qpp
        walls = []
        for rep in range(2):
            s0 = time.perf_counter()
            a0 = [scanner.scan_keys(m, keys[0][i:i + per]) for i in range(0, 3 * n_queries, per)]
            walls.append((time.perf_counter() - s0) * 1e3)
        a0 = np.concatenate(a0)
        a1 = np.concatenate([scanner.scan_keys(m, keys[1][i:i + per])
                             for i in range(0, 3 * n_queries, per)])
        for q, tg in enumerate(targets):
            exact &= bool(np.array_equal(a0[q] ^ a1[q], oracle.fill_rows(seed, tg, 1, w)[0]))
        bm = statistics.median(walls)
        batched.append({"queries_per_pass": qpp, "queries": n_queries, "batch_ms": bm,
                        "queries_per_s": n_queries * 1e3 / bm,
                        "vs_single_query": (n_queries * 1e3 / bm) / (1e3 / pm)})
    out["batched_passes"] = batched
    out["bit_exact"] = exact
    m.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rows-log2", type=int, default=31)
    ap.add_argument("--rows", type=int, default=0, help="explicit row count (overrides --rows-log2)")
    ap.add_argument("--kernel", choices=["auto", "lane_dfs", "stream", "point", "suffix", "split"],
                    default="auto")
    ap.add_argument("--ref-plain-steps", type=int, default=1,
                    help="passes of the reference-shaped per-row suffix-walk kernel, 3 keys each")
    ap.add_argument("--keys-per-wave", type=int, default=0)
    ap.add_argument("--waves-per-block", type=int, default=0)
    ap.add_argument("--row-bytes", type=int, default=32)
    ap.add_argument("--domain-bits", type=int, default=0, help="0 = rows-log2")
    ap.add_argument("--keys", type=int, default=24,
                    help="keys per batched pass: two-party pairs, plus one single key when odd")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--lane-depth", type=int, default=0)
    ap.add_argument("--waves-per-key", type=int, default=0)
    ap.add_argument("--use-graph", action="store_true",
                    help="time the headline steps from a HIP graph")
    ap.add_argument("--cpu-sample-log2", type=int, default=27)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=31)
    ap.add_argument("--q-sweep", default="3,96,384",
                    help="comma-separated key counts: one extra timed pass each, reported in by_keys")
    ap.add_argument("--delta-rows", type=int, default=1 << 16,
                    help="epoch DeltaBuffer entries applied in every pass (uniform rows, random "
                         "diffs); 0 = none")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other workloads (by_config)")
    ap.add_argument("--serve-seconds", type=float, default=20.0,
                    help="seconds of Poisson arrivals in the bounded serving run (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=5.0)
    ap.add_argument("--serve-deadline-ms", type=int, default=600)
    ap.add_argument("--tuned", action="store_true",
                    help="time the autotuner's stored setting for this shape "
                         "(morphogenesis_amd/tuned.json) when there is one")
    ap.add_argument("--autotune", action="store_true",
                    help="tune this shape first (every candidate checked bit for bit against "
                         "the first) and time the winner")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy a 4 GiB matrix once (known bytes for counter scaling)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; the scan shards by row "
                                   "range across GPUs by design but no multi-GPU path is built",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np

    import morphogenesis_amd as mg

    t_bench = time.perf_counter()
    n_rows = args.rows or (1 << args.rows_log2)
    w = args.row_bytes
    d = args.domain_bits or max(1, (n_rows - 1).bit_length())
    q = max(1, args.keys)
    scanner = mg.GpuScanner(0)
    info = scanner.info()
    log(f"device {info['name']} CUs={info['compute_units']} free={info['free_bytes'] / 2**30:.1f} GiB")

    if args.calibrate_copy:
        ca = mg.GpuPageMatrix(scanner, 1 << 27, 32)
        cb = mg.GpuPageMatrix(scanner, 1 << 27, 32)
        ca.fill_seeded(1)
        cb.copy_from(ca)
        log("calibration copy: 4 GiB read + 4 GiB written by copy_u4")
        ca.close()
        cb.close()

    t0 = time.perf_counter()
    m = mg.GpuPageMatrix(scanner, n_rows, w)
    m.fill_seeded(args.seed)
    log(f"matrix {n_rows} x {w} B ({n_rows * w / 2**30:.1f} GiB) filled in "
        f"{time.perf_counter() - t0:.1f} s")

    rng = np.random.default_rng(args.seed)
    # q keys: q // 2 two-party pairs, and with odd q one more key whose
    # partner is scanned once, untimed, to check it.
    targets = [int(x) for x in rng.integers(0, n_rows, size=(q + 1) // 2)]
    pairs = [mg.generate_keys(d, t, seed=1000 + i) for i, t in enumerate(targets)]
    keys = [k for p in pairs for k in p][:q]
    # The epoch DeltaBuffer every pass applies on top of the matrix: uniform
    # rows with replacement, random diffs; the first target is forced into it
    # so the bit-exact check covers the delta path. Answers are for the
    # matrix XOR the delta (scan_delta_for_gpu semantics).
    delta = None
    d_idx = np.zeros(0, dtype=np.uint64)
    d_diffs = np.zeros((0, w), dtype=np.uint8)
    if args.delta_rows:
        d_idx = rng.integers(0, n_rows, size=args.delta_rows).astype(np.uint64)
        d_idx[0] = targets[0]
        d_diffs = rng.integers(0, 256, size=(args.delta_rows, w), dtype=np.uint8)
        delta = mg.DeltaBuffer(scanner, w, args.delta_rows)
        delta.set(d_idx, d_diffs, epoch=1)
        log(f"delta: {args.delta_rows} rows applied in every pass")

    def want_row(t):
        """Row t of the matrix XOR every delta diff aimed at it."""
        row = m.download(t, 1)[0].copy()
        for j in np.nonzero(d_idx == np.uint64(t))[0]:
            row ^= d_diffs[j]
        return row

    target_rows = [want_row(t) for t in targets]
    kernel = {"auto": mg.KERNEL_AUTO, "lane_dfs": mg.KERNEL_LANE_DFS,
              "stream": mg.KERNEL_STREAM, "point": mg.KERNEL_POINT,
              "suffix": mg.KERNEL_SUFFIX, "split": mg.KERNEL_STREAM_SPLIT}[args.kernel]
    tuned = None
    if args.autotune or args.tuned:
        from morphogenesis_amd import autotune
        if args.autotune:
            best, _ = autotune.tune(scanner, m, keys, reps=2, log=log)
            tuned = {k: best[k] for k in autotune.SETTING_KEYS if k in best} if best else None
        else:
            tuned = autotune.lookup(autotune.device_arch(scanner), w, n_rows, q)
        log(f"tuned setting: {tuned}")

    def opts_for(use_graph):
        if tuned:
            return mg.make_opts(**tuned, use_graph=use_graph, delta=delta)
        return mg.make_opts(kernel=kernel, lane_depth=args.lane_depth,
                            keys_per_wave=args.keys_per_wave, waves_per_key=args.waves_per_key,
                            waves_per_block=args.waves_per_block, use_graph=use_graph,
                            delta=delta)

    opts = opts_for(args.use_graph)
    odd_partner = scanner.scan_keys(m, [pairs[-1][1]], opts)[0] if q % 2 else None

    def check(ans):
        for i in range(q // 2):
            if not np.array_equal(ans[2 * i] ^ ans[2 * i + 1], target_rows[i]):
                raise SystemExit(f"bit-exactness failure: pair {i} target {targets[i]}")
        if q % 2 and not np.array_equal(ans[q - 1] ^ odd_partner, target_rows[-1]):
            raise SystemExit(f"bit-exactness failure: single key, target {targets[-1]}")

    for i in range(args.warmup):
        ans, t = scanner.scan_keys(m, keys, opts, with_timing=True)
        check(ans)
        log(f"warmup {i}: {t['total_ms']:.1f} ms (kernel {t['kernel']}, kl={t['lane_depth']}, "
            f"waves={t['waves']})")

    step_ms, scan_ms = [], []
    t_start = time.perf_counter()
    for i in range(args.steps):
        s0 = time.perf_counter()
        ans, t = scanner.scan_keys(m, keys, opts, with_timing=True)
        step_ms.append((time.perf_counter() - s0) * 1e3)
        scan_ms.append(t["scan_kernel_ms"])
        check(ans)
        log(f"step {i}: {step_ms[-1]:.1f} ms, fused kernel {t['scan_kernel_ms']:.1f} ms")
    wall = time.perf_counter() - t_start
    timing = t

    # The same pass from a HIP graph (or eagerly, when the steps used the graph).
    alt = opts_for(not args.use_graph)
    alt_ms, alt_kern = [], []
    for i in range(3):
        s0 = time.perf_counter()
        ans, t_alt = scanner.scan_keys(m, keys, alt, with_timing=True)
        check(ans)
        if i:
            alt_ms.append((time.perf_counter() - s0) * 1e3)
            alt_kern.append(t_alt["scan_kernel_ms"])
    graph = {"steps_use_graph": bool(args.use_graph),
             "eager_pass_ms": statistics.mean(alt_ms if args.use_graph else step_ms),
             "graph_pass_ms": statistics.mean(step_ms if args.use_graph else alt_ms),
             "eager_kernel_ms": statistics.mean(alt_kern if args.use_graph else scan_ms),
             "graph_kernel_ms": statistics.mean(scan_ms if args.use_graph else alt_kern),
             "bit_exact": True}
    log(f"graph replay: {graph['graph_pass_ms']:.1f} ms against eager {graph['eager_pass_ms']:.1f} ms")

    # Plain path: one query (3 keys) per pass over the whole matrix.
    plain_ms = []
    plain_keys = (keys * KEYS_PER_QUERY)[:KEYS_PER_QUERY]
    for i in range(args.plain_steps):
        s0 = time.perf_counter()
        scanner.scan_keys(m, plain_keys, opts)
        plain_ms.append((time.perf_counter() - s0) * 1e3)
        log(f"plain pass {i}: {plain_ms[-1]:.1f} ms")

    # Reference-shaped plain kernel: one query per pass, an 11-level path walk
    # per row from a per-tile prefix (fused_kernel.cu:309-360), same matrix.
    ref_ms = []
    if w <= 128:
        ref_opts = mg.make_opts(kernel=mg.KERNEL_SUFFIX, delta=delta)
        for i in range(args.ref_plain_steps):
            s0 = time.perf_counter()
            ans_ref = scanner.scan_keys(m, list(pairs[0]), ref_opts)
            ref_ms.append((time.perf_counter() - s0) * 1e3 * KEYS_PER_QUERY / 2)
            if not np.array_equal(ans_ref[0] ^ ans_ref[1], target_rows[0]):
                raise SystemExit("bit-exactness failure in the suffix-walk baseline")
            log(f"reference-shaped plain pass {i}: {ref_ms[-1]:.1f} ms per 3 keys")

    # Single passes at other batch sizes (2^31 x 32 B: Q in {3, 24, 96, 384}).
    by_keys = {}
    for qs in [int(x) for x in args.q_sweep.split(",") if x.strip()]:
        qs = max(1, qs)
        sw_targets = [int(x) for x in rng.integers(0, n_rows, size=(qs + 1) // 2)]
        sw_keys = [k for i, t in enumerate(sw_targets)
                   for k in mg.generate_keys(d, t, seed=50_000 + i)][:qs]
        s0 = time.perf_counter()
        ans, t = scanner.scan_keys(m, sw_keys, opts, with_timing=True)
        wall_q = time.perf_counter() - s0
        exact = all(np.array_equal(ans[2 * i] ^ ans[2 * i + 1], want_row(tt))
                    for i, tt in enumerate(sw_targets[:qs // 2]))
        if not exact:
            raise SystemExit(f"bit-exactness failure in the {qs}-key pass")
        by_keys[qs] = {"pass_ms": wall_q * 1e3, "kernel_ms": t["scan_kernel_ms"],
                       "queries_per_s": qs / KEYS_PER_QUERY / wall_q,
                       "valu_frac": valu_frac(n_rows, w, qs, t["scan_kernel_ms"]),
                       "kernel": t["kernel"], "lane_depth": t["lane_depth"],
                       "waves": t["waves"], "pairs_checked": qs // 2, "bit_exact": exact}
        log(f"{qs}-key pass: {wall_q * 1e3:.1f} ms ({by_keys[qs]['queries_per_s']:.2f} queries/s)")

    queries_per_step = q / KEYS_PER_QUERY
    qps = queries_per_step * args.steps / wall
    plain_qps = 1e3 / statistics.mean(plain_ms) if plain_ms else None
    avg_scan_ms = statistics.mean(scan_ms)
    leaves = n_rows * q
    alg_ops = leaves * (OPS_PER_NODE + w // 4)
    valu_tops = alg_ops / (avg_scan_ms * 1e-3) / 1e12
    scan_gbs = n_rows * w / (avg_scan_ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        try:
            with open(tfile) as f:
                traffic = json.load(f).get(f"{n_rows}x{w}x{q}")
        except Exception:
            traffic = None

    # The bound is whichever roofline allows the longer time for this shape:
    # integer VALU for 32-byte rows, HBM for wide rows at a few keys per pass.
    valu_side = {"achieved": valu_tops, "peak": VALU_PEAK_TOPS, "unit": "TOP/s",
                 "frac": valu_tops / VALU_PEAK_TOPS, "ops_per_leaf": OPS_PER_NODE + w // 4,
                 "leaves_per_launch": leaves,
                 "mix_peak": VALU_MIX_PEAK_TOPS, "mix_frac": valu_tops / VALU_MIX_PEAK_TOPS}
    hbm_side = {"achieved": scan_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": scan_gbs / HBM_PEAK_GBS, "bytes_per_launch": n_rows * w}
    if alg_ops / (VALU_PEAK_TOPS * 1e12) >= n_rows * w / (HBM_PEAK_GBS * 1e9):
        roofline = dict(bound="valu", **valu_side, traffic=traffic, kernel_ms=avg_scan_ms,
                        hbm=hbm_side)
    else:
        roofline = dict(bound="hbm", **hbm_side, traffic=traffic, kernel_ms=avg_scan_ms,
                        valu=valu_side)

    by_config = {}
    if not args.no_configs and args.serve_seconds > 0:
        # Serving on the same matrix (its rows are the seeded ones: the delta
        # above is a separate buffer the loop never sees).
        from morphogenesis_amd import serve_bench
        log(f"serving: {args.serve_seconds} s of Poisson arrivals at {args.serve_rate} req/s, "
            f"deadline {args.serve_deadline_ms} ms")
        srv = serve_bench.run(scanner, m, args.seed, rate=args.serve_rate,
                              duration=args.serve_seconds, drain=30.0,
                              deadline_ms=args.serve_deadline_ms, log=log)
        srv_stats = srv.pop("server")
        srv["sweep_steps"] = srv_stats["sweep_steps"]
        srv["model_key_step_ms"] = srv_stats["key_step_ms"]
        srv["model_step_fixed_ms"] = srv_stats["step_fixed_ms"]
        srv["vs_plain"] = (srv["queries_per_s"] / srv["plain_queries_per_s"]
                           if srv["queries_per_s"] and srv["plain_queries_per_s"] else None)
        by_config["serving_2e31x32"] = srv
        log(f"serving: {srv['completed']} done, {srv['refused']} refused, p50 {srv['p50_ms']} "
            f"p99 {srv['p99_ms']} ms, exact {srv['bit_exact']}")

    rows_label = f"2^{n_rows.bit_length() - 1}" if n_rows & (n_rows - 1) == 0 else str(n_rows)
    # Objects go before the device that holds them (closing it first is refused).
    if delta is not None:
        delta.close()
    m.close()

    if not args.no_configs:
        by_config["plain_cpu_2e20x32"] = config_small(
            mg, scanner, 1 << 20, 32, 20, 2020, "2^20 rows x 32 B, one key pair, 20-bit domain")
        log(f"2^20: {by_config['plain_cpu_2e20x32']['one_key'][0]['kernel_ms']:.3f} ms per key")
        by_config["plain_gpu_2e27x32"] = config_small(
            mg, scanner, 1 << 27, 32, 27, 2727, "2^27 rows x 32 B, one key pair, 27-bit domain")
        c2 = by_config["plain_gpu_2e27x32"]
        log(f"2^27: one key {c2['one_key'][0]['kernel_ms']:.2f} ms, pair "
            f"{c2['pair_one_pass']['kernel_ms']:.2f} ms (VALU {c2['pair_one_pass']['valu_frac']:.3f})")
        by_config["pages_78643200x1024"] = config_pages(mg, scanner, 7864)
        for label, n_pages, d_pages, ref in REFERENCE_PAGE_SHAPES:
            by_config[label] = config_reference_pages(mg, scanner, n_pages, d_pages, ref,
                                                      seed=n_pages % 1000 + 1)
            c = by_config[label]
            best = max(c["batched_passes"], key=lambda b: b["queries_per_s"])
            log(f"{label}: single query {c['single_query']['pass_ms']:.3f} ms, "
                f"{c['per_query_streams']['queries_per_s']:.1f} q/s over streams, "
                f"{best['queries_per_s']:.1f} q/s batched {best['queries_per_pass']} per pass, "
                f"exact {c['bit_exact']}")

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        # Sample at most 4 GiB of host rows.
        sample_log2 = min(args.cpu_sample_log2, n_rows.bit_length() - 1, 32 - (w - 1).bit_length())
        cpu = cpu_baseline(n_rows, w, d, sample_log2, args.seed)
        log(f"cpu baseline {cpu['value']:.4f} queries/s on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic: splitmix64-seeded uniform rows generated in HBM, seeded DPF keygen, "
                "uniform targets",
        "config": {"workload": f"{rows_label} rows x {w} B + a {args.delta_rows}-row epoch "
                               f"delta, {q} DPF keys per pass over a {d}-bit domain "
                               f"({q // 2} two-party pairs"
                               f"{' and 1 single key' if q % 2 else ''})",
                   "rows": n_rows, "row_bytes": w, "domain_bits": d, "keys_per_pass": q,
                   "delta_rows": args.delta_rows,
                   "kernel": timing["kernel"], "lane_depth": timing["lane_depth"],
                   "waves": timing["waves"], "use_graph": bool(args.use_graph),
                   "tuned": tuned},
        "p50_ms": percentile(step_ms, 50),
        "p99_ms": percentile(step_ms, 99),
        "keys_per_s": q * args.steps / wall,
        "plain_queries_per_s": plain_qps,
        "batched_vs_plain": (qps / plain_qps) if plain_qps else None,
        "plain_reference_shaped_queries_per_s": (1e3 / statistics.mean(ref_ms)) if ref_ms else None,
        "batched_vs_reference_shaped_plain": (qps * statistics.mean(ref_ms) / 1e3) if ref_ms else None,
        "scan_gbs": scan_gbs,
        "by_keys": by_keys or None,
        "graph": graph,
        "by_config": by_config or None,
        "bit_exact": True,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "bench_seconds": time.perf_counter() - t_bench,
    }
    print(json.dumps(out), flush=True)
    scanner.close()


if __name__ == "__main__":
    main()
