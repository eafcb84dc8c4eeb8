"""GPU parity: the HIP scan against the CPU oracle, byte for byte.

Runs through the C ABI (morphogenesis_amd -> libmorphogen_hip.so). The oracle
only supplies expected answers. Covers the reference's own test cases
(kernel.rs:1443-1677, scan.rs:1274-1572, network_api.rs batch == singles) at
W-byte rows, plus every kernel variant, tile/range edges, fuzzed control
bytes, delta scans and the device fills.
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
    print("device:", s.info(), flush=True)
    yield s
    s.close()


def upload(mg, scanner, rows):
    return mg.GpuPageMatrix.new(scanner, rows)


def pairs_for(d, targets, seed=0):
    return [oracle.keygen_bytes(d, t, seed=seed + 97 * i + t) for i, t in enumerate(targets)]


SHAPES = [
    # n_rows, row_bytes, domain bits
    (1, 32, 1), (2, 16, 1), (3, 32, 2), (255, 32, 8), (256, 48, 8), (500, 64, 10),
    (2047, 32, 11), (2048, 32, 11), (2049, 32, 12), (4096, 128, 12), (5000, 80, 13),
    (65536, 32, 16), (100_003, 32, 20), (1 << 20, 32, 20),
]


@pytest.mark.parametrize("n,w,d", SHAPES)
def test_lane_dfs_matches_oracle_and_recovers(mg, scanner, n, w, d):
    print(f"case n={n} w={w} d={d}", flush=True)
    rows = oracle.fill_rows(seed=n + w, first_row=0, n_rows=n, row_bytes=w)
    m = upload(mg, scanner, rows)
    targets = sorted({0, n - 1, n // 2, (n * 7) // 9})
    pairs = pairs_for(d, targets, seed=d)
    keys0 = [p[0] for p in pairs]
    keys1 = [p[1] for p in pairs]
    g0, t = scanner.scan_keys(m, keys0, with_timing=True)
    assert t["kernel"] == "lane_dfs"
    g1 = scanner.scan_keys(m, keys1)
    assert np.array_equal(g0, oracle.scan_tiled(keys0, rows, THREADS))
    assert np.array_equal(g1, oracle.scan_tiled(keys1, rows, THREADS))
    for q, tgt in enumerate(targets):
        assert np.array_equal(g0[q] ^ g1[q], rows[tgt])
    m.close()


@pytest.mark.parametrize("kl", [1, 2, 3, 5, 8, 10, 12])
def test_lane_depths_agree(mg, scanner, kl):
    n, w, d = 30_000, 32, 15
    rows = oracle.fill_rows(seed=9, first_row=0, n_rows=n, row_bytes=w)
    m = upload(mg, scanner, rows)
    keys = [p[i % 2] for i, p in enumerate(pairs_for(d, [1, 20_000, 29_999], seed=kl))]
    ref = oracle.scan_tiled(keys, rows, THREADS)
    for wpk in (0, 1, 7):
        got = scanner.scan_keys(m, keys, mg.make_opts(kernel=mg.KERNEL_LANE_DFS, lane_depth=kl,
                                                     waves_per_key=wpk))
        assert np.array_equal(got, ref), (kl, wpk)
    # The top 0-4 levels of the lanes' DFS stacks in global memory, the rest in
    # LDS (default and seven waves per key, so waves share blocks and sit
    # idle), with the rows loaded before and after each step's expansions.
    for gs in range(0, min(4, kl - 1) + 1):
        for wpk in (0, 7):
            for late in (False, True):
                got, t = scanner.scan_keys(m, keys, mg.make_opts(
                    kernel=mg.KERNEL_LANE_DFS, lane_depth=kl, waves_per_key=wpk,
                    stack_global=gs, late_loads=late), with_timing=True)
                assert t["stack_global"] == gs
                assert np.array_equal(got, ref), (kl, gs, wpk, late)
    for bad in (kl, 5):
        if bad <= kl - 1 and bad <= 4:
            continue
        with pytest.raises(mg.MgError):
            scanner.scan_keys(m, keys, mg.make_opts(kernel=mg.KERNEL_LANE_DFS, lane_depth=kl,
                                                    stack_global=bad))
    m.close()


def test_tile_root_table_and_prefix_walk_agree(mg, scanner):
    # Launches with global stack levels start tiles from a precomputed table of the
    # keys' tile roots when it fits 16 MiB, else each wave walks (and caches) its tile
    # prefixes. 2^22 rows in 256-row tiles (kl = 2): 8 keys take the table (2.6 MB),
    # 64 keys the walk (21 MB); the first 8 keys' answers must agree, and equal the
    # oracle, over runs that start inside tiles too.
    n, w, d = 1 << 22, 32, 22
    rows = oracle.fill_rows(seed=22, first_row=0, n_rows=n, row_bytes=w)
    m = upload(mg, scanner, rows)
    targets = [int(x) for x in np.random.default_rng(22).integers(0, n, size=32)]
    pairs = pairs_for(d, targets, seed=5)
    keys = [p[0] for p in pairs] + [p[1] for p in pairs]
    ref = oracle.scan_fast(keys[:8], rows, THREADS)
    for wpk in (0, 333):
        opts = mg.make_opts(kernel=mg.KERNEL_LANE_DFS, lane_depth=2, stack_global=1,
                            late_loads=True, waves_per_key=wpk)
        few = scanner.scan_keys(m, keys[:8], opts)
        many = scanner.scan_keys(m, keys, opts)
        assert np.array_equal(few, ref), wpk
        assert np.array_equal(many[:8], ref), wpk
        for q, tgt in enumerate(targets):
            assert np.array_equal(many[q] ^ many[q + 32], rows[tgt]), (wpk, q)
    m.close()


@pytest.mark.parametrize("w", [16, 48, 128])
def test_global_stack_levels_by_row_width(mg, scanner, w):
    # 2^17 rows at kl = 9: every stack split, and the planner's own choice.
    n, d = 1 << 17, 17
    rows = oracle.fill_rows(seed=w, first_row=0, n_rows=n, row_bytes=w)
    m = upload(mg, scanner, rows)
    keys = [p[i % 2] for i, p in enumerate(pairs_for(d, [5, 70_000, n - 1], seed=w))]
    ref = oracle.scan_tiled(keys, rows, THREADS)
    for gs in (None, 0, 1, 3, 4):
        for late in (None, False, True):
            got = scanner.scan_keys(m, keys, mg.make_opts(kernel=mg.KERNEL_LANE_DFS, lane_depth=9,
                                                         stack_global=gs, late_loads=late))
            assert np.array_equal(got, ref), (w, gs, late)
    bad = mg.make_opts(kernel=mg.KERNEL_LANE_DFS)
    bad.row_loads = 3
    with pytest.raises(mg.MgError):
        scanner.scan_keys(m, keys, bad)
    m.close()


# Rows of <= 16 chunks go 4 to a load instruction, of 17-32 chunks 2 to a
# load (144, 256, 16, 32 / 272, 512 B); 528 B rows and wider go one to a load.
@pytest.mark.parametrize("n,w,d", [(1000, 256, 10), (3000, 1024, 12), (777, 4096, 10),
                                   (5000, 2048, 13), (2500, 272, 12), (600, 32, 10),
                                   (3001, 144, 12), (2000, 512, 11), (4000, 16, 12),
                                   (1500, 528, 11), (70_001, 144, 17), (40_000, 272, 16)])
def test_stream_kernel_matches_oracle(mg, scanner, n, w, d):
    print(f"case n={n} w={w} d={d}", flush=True)
    rows = oracle.fill_rows(seed=n * 3 + w, first_row=0, n_rows=n, row_bytes=w)
    m = upload(mg, scanner, rows)
    targets = [0, n - 1, n // 3, n // 2]
    pairs = pairs_for(d, targets, seed=w)
    keys = [p[0] for p in pairs]
    ref = oracle.scan_tiled(keys, rows, THREADS)
    max_g = 4 if w <= 1024 else (3 if w <= 2048 else 2)
    for g in range(1, max_g + 1):
        for kl, wpb in ((1, 1), (4, 2), (6, 1), (7, 1)):
            got, t = scanner.scan_keys(m, keys, mg.make_opts(kernel=mg.KERNEL_STREAM, lane_depth=kl,
                                                            keys_per_wave=g, waves_per_block=wpb),
                                       with_timing=True)
            assert t["kernel"] == "stream"
            assert np.array_equal(got, ref), (g, kl, wpb)
    k1 = [p[1] for p in pairs]
    g1 = scanner.scan_keys(m, k1)
    for q, tgt in enumerate(targets):
        assert np.array_equal(ref[q] ^ g1[q], rows[tgt])
    m.close()


def split_pair_bytes(g, kl):
    """LDS of one split-kernel pair at the smallest mask ring: g stacks of the kl - 1
    DFS levels less those held in registers (three, five for groups of four keys;
    64 x 16 B each), two 128-mask buffers per key and 16 B of hand-over counters. A
    block gets 160 KiB of LDS."""
    reg = 5 if g >= 4 else 3
    return (g * max(0, kl - 1 - reg) * 64 + 256 * g + 1) * 16


@pytest.mark.parametrize("n,w,d", [(3000, 1024, 12), (5000, 2048, 13), (777, 4096, 10),
                                   (2500, 272, 12), (600, 32, 10), (70_001, 1024, 17),
                                   (3001, 3088, 12), (9000, 4096, 14), (2100, 1040, 12),
                                   (1800, 1552, 11)])
def test_split_stream_kernel_matches_oracle(mg, scanner, n, w, d):
    # Producer/consumer blocks: masks from a DFS wave, rows from a streaming wave.
    print(f"case split n={n} w={w} d={d}", flush=True)
    rows = oracle.fill_rows(seed=n * 5 + w, first_row=0, n_rows=n, row_bytes=w)
    m = upload(mg, scanner, rows)
    targets = [0, n - 1, n // 3, n // 2, n // 5]
    keys = [p[i % 2] for i, p in enumerate(pairs_for(d, targets, seed=w + 1))]
    ref = oracle.scan_tiled(keys, rows, THREADS)
    max_g = 4 if w <= 2048 else 3
    # 2, 4, 8 waves per block = 1, 2, 4 producer/consumer pairs; pairs past the
    # work and pairs with fewer tiles than the block's others run empty steps.
    # Lane depths 1-5 cover DFS stacks of no slot, one to three register slots, and
    # three register slots around one LDS level.
    for g in range(1, max_g + 1):
        for kl in sorted(k for k in {1, 2, 3, 4, 5, d - 6} if k <= d - 6):
            for wpb in (0, 4, 8):
                opts = mg.make_opts(kernel=mg.KERNEL_STREAM_SPLIT, lane_depth=kl,
                                    keys_per_wave=g, waves_per_block=wpb)
                if (wpb or 2) // 2 * split_pair_bytes(g, kl) > 160 * 1024:
                    with pytest.raises(mg.MgError):
                        scanner.scan_keys(m, keys, opts)
                    continue
                got, t = scanner.scan_keys(m, keys, opts, with_timing=True)
                assert t["kernel"] == "stream_split"
                assert np.array_equal(got, ref), (g, kl, wpb)
    # Mask rings of 2 to 8 leaf-pair steps (a DFS wave runs up to depth - 1 steps ahead).
    for depth in (2, 3, 5, 8):
        opts = mg.make_opts(kernel=mg.KERNEL_STREAM_SPLIT, lane_depth=d - 6, mask_depth=depth)
        try:
            got = scanner.scan_keys(m, keys, opts)
        except mg.MgError as e:
            assert "LDS" in str(e), e  # a deep ring of many keys can exceed a block's LDS
            continue
        assert np.array_equal(got, ref), depth
    with pytest.raises(mg.MgError):
        scanner.scan_keys(m, keys, mg.make_opts(kernel=mg.KERNEL_STREAM_SPLIT, mask_depth=1))
    # Uneven parts: a waves-per-key count that leaves the last part short.
    for wpb in (2, 8):
        got = scanner.scan_keys(m, keys, mg.make_opts(kernel=mg.KERNEL_STREAM_SPLIT, lane_depth=1,
                                                     waves_per_key=7, waves_per_block=wpb))
        assert np.array_equal(got, ref), wpb
    # Pairs take runs of leaf-pair steps that start and end inside a tile, so a
    # producer first walks down to its run's first pair: 3 and 7 runs per key,
    # and 1000 (one or two steps per pair).
    for g in range(1, max_g + 1):
        for kl in sorted({3, d - 6}):
            for wpk in (3, 7, 1000):
                for wpb in (2, 8):
                    if wpb // 2 * split_pair_bytes(g, kl) > 160 * 1024:
                        continue
                    opts = mg.make_opts(kernel=mg.KERNEL_STREAM_SPLIT, lane_depth=kl,
                                        keys_per_wave=g, waves_per_key=wpk, waves_per_block=wpb)
                    got = scanner.scan_keys(m, keys, opts)
                    assert np.array_equal(got, ref), (g, kl, wpk, wpb)
    # Rows over 1 KiB in two column halves, one pair each, up to 4 keys per pair. At
    # 1040 and 1552 B the second half is ragged (1 and 33 of its 64 chunks), at 3088 B
    # ragged with two chunks per lane.
    for g in range(1, 5):
        for kl in sorted({3, d - 6}):
            for wpb in (0, 2, 4, 8):
                for wpk in (0, 7):
                    opts = mg.make_opts(kernel=mg.KERNEL_STREAM_SPLIT, lane_depth=kl,
                                        keys_per_wave=g, waves_per_block=wpb, waves_per_key=wpk,
                                        row_halves=True)
                    if w <= 1024:
                        with pytest.raises(mg.MgError):
                            scanner.scan_keys(m, keys, opts)
                        continue
                    if (wpb or 2) // 2 * split_pair_bytes(g, kl) > 160 * 1024:
                        continue
                    got, t = scanner.scan_keys(m, keys, opts, with_timing=True)
                    assert np.array_equal(got, ref), (g, kl, wpb, wpk)
    bad_halves = mg.make_opts(kernel=mg.KERNEL_STREAM_SPLIT)
    bad_halves.row_halves = 3
    with pytest.raises(mg.MgError):
        scanner.scan_keys(m, keys, bad_halves)
    for bad in (1, 3, 6):
        with pytest.raises(mg.MgError):
            scanner.scan_keys(m, keys, mg.make_opts(kernel=mg.KERNEL_STREAM_SPLIT,
                                                    waves_per_block=bad))
    # Too shallow a domain for 64 live lanes.
    with pytest.raises(mg.MgError):
        scanner.scan_keys(m, keys, mg.make_opts(kernel=mg.KERNEL_STREAM_SPLIT, lane_depth=d - 5))
    m.close()


@pytest.mark.parametrize("n,w,d", [(20_000, 2048, 15), (9000, 4096, 14), (7001, 2048, 13)])
def test_paced_key_groups_match_oracle(mg, scanner, n, w, d):
    # Passes of several key groups over rows of two chunks per lane (2 KiB rows, and the
    # column halves of 4 KiB rows) pace the streaming waves of the pairs that read the same
    # rows and load the rows through L2. 5-12 keys make 2-3 groups: with three, some pairs
    # find their partners in the next block and run unpaced. Step runs that start inside
    # tiles (waves_per_key 7 and 333) and a ragged last tile (7001 rows) included.
    print(f"case paced n={n} w={w} d={d}", flush=True)
    rows = oracle.fill_rows(seed=n + 3 * w, first_row=0, n_rows=n, row_bytes=w)
    m = upload(mg, scanner, rows)
    targets = [0, n - 1, n // 7, n // 3, n // 2, (2 * n) // 3]
    pairs = pairs_for(d, targets, seed=w + 5)
    keys = [p[0] for p in pairs] + [p[1] for p in pairs]
    ref = oracle.scan_tiled(keys, rows, THREADS)
    for nk in (5, 8, 12):
        for wpk in (0, 7, 333):
            got, t = scanner.scan_keys(m, keys[:nk], mg.make_opts(waves_per_key=wpk),
                                       with_timing=True)
            assert t["kernel"] == "stream_split"
            assert t["keys_per_wave"] == 4, t  # full groups of 4 keys, the last one padded
            assert np.array_equal(got, ref[:nk]), (nk, wpk)
    got = scanner.scan_keys(m, keys)
    for q, tgt in enumerate(targets):
        assert np.array_equal(got[q] ^ got[q + len(targets)], rows[tgt])
    m.close()


@pytest.mark.parametrize("n,w,d", [(3000, 1024, 12), (70_001, 1024, 17), (2100, 528, 12),
                                   (5000, 1008, 13)])
@pytest.mark.parametrize("n_keys", [2, 4])
def test_split_two_dfs_waves_match_oracle(mg, scanner, n, w, d, n_keys):
    # One group of 2 or 4 keys over rows of up to 64 chunks: each pair has two DFS
    # waves with half the keys each, and the streaming wave waits for both.
    print(f"case split-dw2 n={n} w={w} d={d} keys={n_keys}", flush=True)
    rows = oracle.fill_rows(seed=n * 7 + w, first_row=0, n_rows=n, row_bytes=w)
    m = upload(mg, scanner, rows)
    targets = [0, n - 1, n // 3, n // 2][:n_keys]
    keys = [p[i % 2] for i, p in enumerate(pairs_for(d, targets, seed=w + 3))]
    ref = oracle.scan_tiled(keys, rows, THREADS)
    for kl in sorted({1, 2, 3, 5, d - 6}):
        for wpb in (0, 2, 4, 8):
            for wpk in (0, 3, 1000):  # runs that start and end inside tiles
                opts = mg.make_opts(kernel=mg.KERNEL_STREAM_SPLIT, lane_depth=kl,
                                    waves_per_block=wpb, waves_per_key=wpk)
                if (wpb or 2) // 2 * split_pair_bytes(n_keys, kl) > 160 * 1024:
                    continue
                got, t = scanner.scan_keys(m, keys, opts, with_timing=True)
                assert t["kernel"] == "stream_split" and t["dfs_waves"] == 2, t
                assert np.array_equal(got, ref), (kl, wpb, wpk)
    for depth in (2, 3, 4):
        try:
            got = scanner.scan_keys(m, keys, mg.make_opts(kernel=mg.KERNEL_STREAM_SPLIT,
                                                          lane_depth=d - 6, mask_depth=depth))
        except mg.MgError as e:
            assert "LDS" in str(e), e
            continue
        assert np.array_equal(got, ref), depth
    # Several groups keep one DFS wave per pair.
    more = keys + [pairs_for(d, [n // 7], seed=w + 9)[0][0]] * 5
    got, t = scanner.scan_keys(m, more, None, with_timing=True)
    assert t["kernel"] == "stream_split" and t["dfs_waves"] == 1, t
    assert np.array_equal(got, oracle.scan_tiled(more, rows, THREADS))
    m.close()


# Kernel AUTO picks: lane-DFS up to 128-byte rows; the split kernel for
# 528-1008-byte rows with >= 3 keys and 1024-4096-byte rows, when d >= 11
# (rows over 2 KiB in column halves from 4 keys); the stream kernel otherwise. Small domains shrink the split kernel's lane depth.
@pytest.mark.parametrize("n,w,d,q,kernel", [
    (3000, 1024, 12, 1, "stream_split"), (2000, 1024, 11, 1, "stream_split"),
    (1000, 1024, 10, 1, "stream"), (3000, 768, 12, 3, "stream_split"),
    (3000, 768, 12, 2, "stream"), (3000, 512, 12, 3, "stream"), (3000, 128, 12, 3, "lane_dfs"),
    (3000, 2048, 12, 3, "stream_split"), (3000, 2048, 12, 2, "stream_split"),
    (3000, 4096, 12, 3, "stream_split"), (3000, 4096, 12, 1, "stream_split"),
    (1000, 4096, 10, 3, "stream"), (9000, 4096, 14, 4, "stream_split"),
    (3000, 3088, 12, 5, "stream_split"), (9000, 1024, 14, 3, "stream_split"),
    (9000, 1024, 14, 4, "stream_split"), (9000, 1024, 14, 8, "stream_split")])
def test_auto_kernel_choice(mg, scanner, n, w, d, q, kernel):
    print(f"case auto n={n} w={w} d={d} q={q}", flush=True)
    rows = oracle.fill_rows(seed=n + w + q, first_row=0, n_rows=n, row_bytes=w)
    m = upload(mg, scanner, rows)
    keys = [p[i % 2] for i, p in enumerate(pairs_for(d, [(n * (i + 1)) // (q + 1)
                                                         for i in range(q)], seed=d))]
    ref = oracle.scan_tiled(keys, rows, THREADS)
    got, t = scanner.scan_keys(m, keys, with_timing=True)
    assert t["kernel"] == kernel
    assert np.array_equal(got, ref)
    # A block size given for another kernel does not stop AUTO's choice.
    got = scanner.scan_keys(m, keys, mg.make_opts(waves_per_block=1))
    assert np.array_equal(got, ref)
    m.close()


def test_point_kernel_and_zero_domain(mg, scanner):
    rows = oracle.fill_rows(seed=4, first_row=0, n_rows=700, row_bytes=48)
    m = upload(mg, scanner, rows)
    keys = [p[0] for p in pairs_for(10, [3, 699])]
    got = scanner.scan_keys(m, keys, mg.make_opts(kernel=mg.KERNEL_POINT))
    assert np.array_equal(got, oracle.scan_tiled(keys, rows, THREADS))
    m.close()
    one = oracle.fill_rows(seed=5, first_row=0, n_rows=1, row_bytes=32)
    m1 = upload(mg, scanner, one)
    z0 = bytearray(35)
    z0[0:16] = bytes(range(16))
    z0[19:35] = bytes([0xA5] * 16)
    z1 = bytearray(z0)
    z1[16] = 1
    for k in (bytes(z0), bytes(z1)):
        assert np.array_equal(scanner.scan_keys(m1, [k])[0], oracle.scan_brute(k, one))
    m1.close()


@pytest.mark.parametrize("n,w,d", [(1, 32, 1), (100, 32, 7), (2048, 16, 11), (5000, 64, 13),
                                   (70_000, 32, 20)])
def test_suffix_walk_baseline_matches_oracle(mg, scanner, n, w, d):
    rows = oracle.fill_rows(seed=n + 7, first_row=0, n_rows=n, row_bytes=w)
    m = upload(mg, scanner, rows)
    keys = [p[i % 2] for i, p in enumerate(pairs_for(d, sorted({0, n - 1, n // 2}), seed=3))]
    got, t = scanner.scan_keys(m, keys, mg.make_opts(kernel=mg.KERNEL_SUFFIX), with_timing=True)
    assert t["kernel"] == "suffix"
    assert np.array_equal(got, oracle.scan_tiled(keys, rows, THREADS))
    m.close()


def test_large_domain_small_matrix(mg, scanner):
    # Keys over 31- and 32-bit domains on a matrix far smaller than 2^d.
    for d in (31, 32):
        n = 100_000
        rows = oracle.fill_rows(seed=d, first_row=0, n_rows=n, row_bytes=32)
        m = upload(mg, scanner, rows)
        targets = [0, 99_999, 65_536]
        pairs = pairs_for(d, targets, seed=d)
        g0 = scanner.scan_keys(m, [p[0] for p in pairs])
        g1 = scanner.scan_keys(m, [p[1] for p in pairs])
        assert np.array_equal(g0, oracle.scan_tiled([p[0] for p in pairs], rows, THREADS))
        for q, t in enumerate(targets):
            assert np.array_equal(g0[q] ^ g1[q], rows[t])
        m.close()


def test_fuzzed_control_bytes(mg, scanner):
    rng = np.random.default_rng(11)
    n, w, d = 3000, 32, 12
    rows = oracle.fill_rows(seed=12, first_row=0, n_rows=n, row_bytes=w)
    m = upload(mg, scanner, rows)
    keys = []
    for trial in range(8):
        b = bytearray(oracle.keygen_bytes(d, int(rng.integers(0, n)), seed=trial)[trial % 2])
        b[16] = int(rng.choice([0, 1, 2, 0xFF]))
        for lvl in range(d):
            off = 19 + 18 * lvl
            b[off + 16] = int(rng.choice([0, 1, 3, 0x80]))
            b[off + 17] = int(rng.choice([0, 1, 2, 0xFE]))
        keys.append(bytes(b))
    ref = oracle.scan_tiled(keys, rows, THREADS)
    for kern in (mg.KERNEL_LANE_DFS, mg.KERNEL_STREAM, mg.KERNEL_POINT, mg.KERNEL_SUFFIX):
        assert np.array_equal(scanner.scan_keys(m, keys, mg.make_opts(kernel=kern)), ref), kern
    m.close()


def test_row_ranges_compose(mg, scanner):
    n, w, d = 200_000, 32, 18
    rows = oracle.fill_rows(seed=21, first_row=0, n_rows=n, row_bytes=w)
    m = upload(mg, scanner, rows)
    keys = [p[0] for p in pairs_for(d, [5, 150_000, 199_999])]
    full = scanner.scan_keys(m, keys)
    assert np.array_equal(full, oracle.scan_tiled(keys, rows, THREADS))
    kl = 6
    tile = 1 << (kl + 6)
    cuts = [0, 3 * tile, 17 * tile, 40 * tile, n]
    acc = np.zeros_like(full)
    for a, b in zip(cuts[:-1], cuts[1:]):
        acc ^= scanner.scan_keys(m, keys, mg.make_opts(lane_depth=kl, row_begin=a, row_end=b))
    assert np.array_equal(acc, full)
    with pytest.raises(mg.RowsError):
        scanner.scan_keys(m, keys, mg.make_opts(lane_depth=kl, row_begin=tile + 1))
    m.close()


def test_graph_replay_matches_eager(mg, scanner):
    n, w, d = 50_000, 32, 16
    rows = oracle.fill_rows(seed=31, first_row=0, n_rows=n, row_bytes=w)
    m = upload(mg, scanner, rows)
    pairs = pairs_for(d, [0, 123, 49_999])
    eager = scanner.scan_keys(m, [p[0] for p in pairs])
    for rep in range(4):
        keys = [p[rep % 2] for p in pairs]
        got = scanner.scan_keys(m, keys, mg.make_opts(use_graph=True))
        assert np.array_equal(got, scanner.scan_keys(m, keys)), rep
        if rep % 2 == 0:
            assert np.array_equal(got, eager)
    m.close()


def test_batch_equals_singles_and_reference_shapes(mg, scanner):
    # network_api.rs:1739-1817 batch == singles; kernel.rs tiled == full batch.
    n, w, d = 4096, 4096, 12
    rows = oracle.fill_rows(seed=41, first_row=0, n_rows=n, row_bytes=w)
    m = upload(mg, scanner, rows)
    targets = [[2048, 100, 3000], [0, 4095, 1], [7, 8, 9], [1000, 2000, 3000]]
    qs0, qs1 = [], []
    for i, tq in enumerate(targets):
        ps = pairs_for(d, tq, seed=i * 11)
        qs0.append([p[0] for p in ps])
        qs1.append([p[1] for p in ps])
    batch0 = scanner.scan_batch(m, qs0)
    batch1 = scanner.scan_batch(m, qs1)
    singles = [scanner.scan(m, q) for q in qs0]
    tiled = scanner.scan_batch_optimized_tiled(m, qs0, tile_size=3)
    per_query = scanner.scan_batch_single_query(m, qs0)
    for i, tq in enumerate(targets):
        assert batch0[i].pages == singles[i].pages == tiled[i].pages == per_query[i].pages
        for j, t in enumerate(tq):
            got = np.frombuffer(batch0[i].pages[j], np.uint8) ^ np.frombuffer(batch1[i].pages[j],
                                                                              np.uint8)
            assert np.array_equal(got, rows[t])
        assert batch0[i].verifs == [bytes(16)] * 3
    m.close()


def test_shape_errors(mg, scanner):
    rows = np.zeros((300, 32), dtype=np.uint8)
    m = upload(mg, scanner, rows)
    k8 = oracle.keygen_bytes(8, 1, seed=1)[0]
    with pytest.raises(mg.RowsError):
        scanner.scan_keys(m, [k8])
    k9 = oracle.keygen_bytes(9, 1, seed=1)[0]
    k10 = oracle.keygen_bytes(10, 1, seed=1)[0]
    with pytest.raises(mg.InvalidDomainBits):
        scanner.scan_keys(m, [k9, k10])
    with pytest.raises(mg.InvalidKeyFormat):
        scanner.scan_keys(m, [k9[:-1]])
    with pytest.raises(mg.RowsError):
        mg.GpuPageMatrix(scanner, 10, 24)
    with pytest.raises(mg.RowsError):
        m.update_rows(299, np.zeros((2, 32), dtype=np.uint8))
    assert scanner.scan_keys(m, []).shape == (0, 32)
    m.close()


def test_delta_scan_and_merge(mg, scanner):
    n, w, d = 100_000, 32, 17
    rows = oracle.fill_rows(seed=51, first_row=0, n_rows=n, row_bytes=w)
    m = upload(mg, scanner, rows)
    rng = np.random.default_rng(3)
    idx = rng.integers(0, n, size=4096).astype(np.uint64)
    idx[:7] = idx[7]
    diffs = rng.integers(0, 256, size=(4096, w), dtype=np.uint8)
    delta = mg.DeltaBuffer(scanner, w, 1 << 16)
    delta.set(idx, diffs, epoch=5)
    assert delta.info() == (4096, 5)
    keys = [p[i % 2] for i, p in enumerate(pairs_for(d, [9, int(idx[7]), 99_999]))]
    dref = oracle.scan_delta(keys, idx, diffs)
    assert np.array_equal(scanner.scan_delta(delta, keys), dref)
    merged = rows.copy()
    for i, r in zip(idx, diffs):
        merged[int(i)] ^= r
    mref = oracle.scan_tiled(keys, merged, THREADS)
    with_delta = scanner.scan_keys(m, keys, mg.make_opts(delta=delta))
    assert np.array_equal(with_delta, mref)
    m.apply_delta(delta)
    assert np.array_equal(m.download(), merged)
    assert np.array_equal(scanner.scan_keys(m, keys), mref)
    delta.close()
    m.close()


def test_device_fills_match_oracle(mg, scanner):
    m = mg.GpuPageMatrix(scanner, 3001, 48)
    m.fill_seeded(77)
    assert np.array_equal(m.download(), oracle.fill_rows(77, 0, 3001, 48))
    m.fill_with_pattern(12345)
    assert np.array_equal(m.download().reshape(-1), oracle.lcg_fill(12345, 0, 3001 * 48))
    m.close()


def test_c2_shape_two_party_recovery(mg, scanner):
    # 2^27 x 32 B (4 GiB), d = 27: bit-exact per server vs the oracle, and the
    # two answers XOR to the target row.
    n, w, d = 1 << 27, 32, 27
    m = mg.GpuPageMatrix(scanner, n, w)
    m.fill_seeded(2027)
    target = 98_765_432
    k0, k1 = oracle.keygen_bytes(d, target, seed=27)
    g = scanner.scan_keys(m, [k0, k1])
    assert np.array_equal(g[0] ^ g[1], oracle.fill_rows(2027, target, 1, w)[0])
    assert np.array_equal(m.download(target, 1)[0], oracle.fill_rows(2027, target, 1, w)[0])
    # Oracle on the first 2^24 rows (a key-aligned prefix) for per-server parity.
    sub = 1 << 24
    rows = oracle.fill_rows(2027, 0, sub, w)
    got = scanner.scan_keys(m, [k0, k1], mg.make_opts(row_end=sub))
    assert np.array_equal(got, oracle.scan_tiled([k0, k1], rows, THREADS))
    m.close()


def test_golden_fixtures(mg, scanner):
    import json
    with open(os.path.join(os.path.dirname(__file__), "golden", "scan_fixtures.json")) as f:
        fx = json.load(f)
    for c in fx["scan_cases"]:
        print("fixture", c["n_rows"], c["row_bytes"], c["domain_bits"], flush=True)
        m = mg.GpuPageMatrix(scanner, c["n_rows"], c["row_bytes"])
        m.fill_seeded(c["data_seed"])
        keys = [bytes.fromhex(k) for k in c["keys"]]
        got = scanner.scan_keys(m, keys)
        assert [a.tobytes().hex() for a in got] == c["answers"]
        m.close()
    for c in fx["delta_cases"]:
        m = mg.GpuPageMatrix(scanner, c["n_rows"], c["row_bytes"])
        m.fill_seeded(c["data_seed"])
        keys = [bytes.fromhex(k) for k in c["keys"]]
        diffs = np.frombuffer(bytes.fromhex(c["diffs"]), np.uint8).reshape(-1, c["row_bytes"])
        dl = mg.DeltaBuffer(scanner, c["row_bytes"], len(c["row_idx"]))
        dl.set(np.array(c["row_idx"], dtype=np.uint64), diffs)
        assert [a.tobytes().hex() for a in scanner.scan_delta(dl, keys)] == c["answers"]
        got = scanner.scan_keys(m, keys, mg.make_opts(delta=dl))
        assert [a.tobytes().hex() for a in got] == c["merged_answers"]
        dl.close()
        m.close()
