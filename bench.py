"""Benchmark of the fused ChaCha8-DPF masked-XOR PIR scan on one MI355X.

Workload (default): the full-mainnet shape, 2^31 rows x 32 B (64 GiB) of
seeded uniform bytes resident in HBM plus an epoch DeltaBuffer of 2^16
updated rows (uniform indices, random diffs), DPF keys over a 31-bit domain
with uniform targets. One step = one batched pass of --keys keys over the
whole matrix and the delta. Keys come in party-0/party-1 pairs so every step
also checks, bit for bit, that each pair's answers XOR to the target row of
the updated matrix.

Reported: batched queries/s (a query = 3 cuckoo keys), per-pass p50/p99
latency, the plain one-query-per-pass rate beside it, scan GB/s and the
fraction of the integer-VALU and HBM rooflines for the fused kernel (HIP
event time on the kernel's stream), and a CPU baseline (the oracle's tiled
scan, all allowed host threads) on a bounded sample of the same workload.
"""
import argparse
import json
import os
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def percentile(xs, p):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * p / 100.0
    lo, hi = int(k), min(int(k) + 1, len(xs) - 1)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def cpu_baseline(n_rows, row_bytes, domain_bits, sample_rows_log2, seed):
    """Oracle tiled scan (C restatement of eval_fused_3dpf_cpu) on a prefix
    of the same matrix; throughput scaled to the full row count."""
    import numpy as np

    import oracle
    threads = max(1, min(16, len(os.sched_getaffinity(0)),
                         int(os.environ.get("OMP_NUM_THREADS", "16") or 16)))
    n = 1 << sample_rows_log2
    rows = oracle.fill_rows(seed, 0, n, row_bytes)
    keys = [oracle.keygen_bytes(domain_bits, (i * 2654435761) % n, seed=100 + i)[0]
            for i in range(KEYS_PER_QUERY)]
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        oracle.scan_tiled(keys, rows, threads)
        times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    full_scale = n_rows / n
    q_per_s = 1.0 / (t * full_scale)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
    except Exception:
        pass
    return {"value": q_per_s, "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"oracle scan_tiled, 3 keys (1 query) over the first 2^{sample_rows_log2} of "
                      f"{n_rows} rows x {row_bytes} B, median of 5 ({t:.2f} s each), scaled "
                      f"x{full_scale:.1f} to the full matrix; cpu: {model}"}


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
    ap.add_argument("--use-graph", action="store_true")
    ap.add_argument("--cpu-sample-log2", type=int, default=27)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=31)
    ap.add_argument("--q-sweep", default="",
                    help="comma-separated key counts: one extra timed pass each, reported in by_keys")
    ap.add_argument("--delta-rows", type=int, default=1 << 16,
                    help="epoch DeltaBuffer entries applied in every pass (uniform rows, random "
                         "diffs); 0 = none")
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
    if tuned:
        opts = mg.make_opts(**tuned, use_graph=args.use_graph, delta=delta)
    else:
        opts = mg.make_opts(kernel=kernel, lane_depth=args.lane_depth,
                            keys_per_wave=args.keys_per_wave, waves_per_key=args.waves_per_key,
                            waves_per_block=args.waves_per_block, use_graph=args.use_graph,
                            delta=delta)
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

    # Optional single passes at other batch sizes (2^31 x 32 B: Q in {3, 24, 96, 384}).
    by_keys = {}
    for qs in [int(x) for x in args.q_sweep.split(",") if x.strip()]:
        qs = max(2, qs - qs % 2)
        sw_targets = [int(x) for x in rng.integers(0, n_rows, size=qs // 2)]
        sw_keys = [k for i, t in enumerate(sw_targets)
                   for k in mg.generate_keys(d, t, seed=50_000 + i)]
        s0 = time.perf_counter()
        ans, t = scanner.scan_keys(m, sw_keys, opts, with_timing=True)
        wall_q = time.perf_counter() - s0
        for i, tt in enumerate(sw_targets[:4]):
            if not np.array_equal(ans[2 * i] ^ ans[2 * i + 1], want_row(tt)):
                raise SystemExit(f"bit-exactness failure in the {qs}-key pass")
        by_keys[qs] = {"pass_ms": wall_q * 1e3, "kernel_ms": t["scan_kernel_ms"],
                       "queries_per_s": qs / KEYS_PER_QUERY / wall_q,
                       "valu_frac": qs * n_rows * (OPS_PER_NODE + w // 4)
                       / (t["scan_kernel_ms"] * 1e-3) / 1e12 / VALU_PEAK_TOPS}
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

    rows_label = f"2^{n_rows.bit_length() - 1}" if n_rows & (n_rows - 1) == 0 else str(n_rows)
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
        "bit_exact": True,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    # Objects go before the device that holds them (closing it first is refused).
    if delta is not None:
        delta.close()
    m.close()
    scanner.close()


if __name__ == "__main__":
    main()
