"""Pins the CPU oracle: OpenSSL ChaCha20, closed forms and DPF invariants.

These mirror the reference's own inline tests (chacha_prg.rs:186-263,
dpf.rs:434-609, kernel.rs:1385-1562, scan.rs:1274-1572) and add what those
leave open: an independent ChaCha implementation (OpenSSL) for the block and
PRG wiring, closed forms, and brute-force cross-checks.
"""
import json
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def words_le(b: bytes):
    return np.frombuffer(b, dtype="<u4").astype(np.uint32)


def ones_mask():
    return np.full(4, 0xFFFFFFFF, dtype=np.uint32)


@pytest.fixture(scope="module")
def openssl_vectors():
    with open(os.path.join(GOLDEN, "chacha20_openssl.json")) as f:
        return json.load(f)


# --------------------------------------------------------------- ChaCha ----

def test_chacha20_block_matches_openssl(openssl_vectors):
    for case in openssl_vectors["block_cases"]:
        key = words_le(bytes.fromhex(case["key"]))
        nonce = words_le(bytes.fromhex(case["nonce"]))
        ks = bytes.fromhex(case["keystream"])
        for blk in range(2):
            out = oracle.chacha_block(key, case["counter"] + blk, nonce, rounds=20)
            assert out.astype("<u4").tobytes() == ks[64 * blk:64 * blk + 64], case


def test_chacha_zero_rounds_is_doubled_state():
    # With no rounds the output is the feed-forward of the initial state onto
    # itself: 2 * (constants | key | counter | nonce) mod 2^32. Pins the state
    # layout independently of the round function.
    key = np.arange(1, 9, dtype=np.uint32) * 0x01010101
    nonce = np.array([0xA, 0xB, 0xC], dtype=np.uint32)
    out = oracle.chacha_block(key, 77, nonce, rounds=0)
    init = np.array([0x61707865, 0x3320646E, 0x79622D32, 0x6B206574, *key, 77, *nonce],
                    dtype=np.uint64)
    assert np.array_equal(out, ((2 * init) & 0xFFFFFFFF).astype(np.uint32))


def test_chacha8_is_a_prefix_of_the_round_sequence():
    # 8 rounds and 20 rounds share the state layout but must differ; 8 rounds
    # must also differ from 6 and 10 (the loop runs rounds/2 double rounds).
    key = np.arange(8, dtype=np.uint32)
    outs = {r: oracle.chacha_block(key, 0, rounds=r).tobytes() for r in (6, 8, 10, 20)}
    assert len(set(outs.values())) == 4


def test_prg_expand_wiring_matches_openssl(openssl_vectors):
    # key = seed || seed, nonce 0: block(counter 0) = left || right seeds,
    # block(counter 1) words 0/1 give the control bits (chacha_prg.rs:153-178).
    for case in openssl_vectors["expand_cases"]:
        seed = words_le(bytes.fromhex(case["seed"]))
        ks = bytes.fromhex(case["keystream"])
        left, right, lt, rt = oracle.prg_expand(seed, rounds=20)
        assert left.astype("<u4").tobytes() == ks[0:16]
        assert right.astype("<u4").tobytes() == ks[16:32]
        assert lt == ks[64] & 1
        assert rt == ks[68] & 1


def test_prg_expand_rounds8_properties():
    # chacha_prg.rs:210-251: deterministic, distinct children, binary t.
    seen = set()
    for i in range(200):
        seed = np.array([i, i * 7 + 1, 0xDEADBEEF ^ i, 0x12345678], dtype=np.uint32)
        l1, r1, lt, rt = oracle.prg_expand(seed)
        l2, r2, lt2, rt2 = oracle.prg_expand(seed)
        assert np.array_equal(l1, l2) and np.array_equal(r1, r2) and (lt, rt) == (lt2, rt2)
        assert not np.array_equal(l1, r1)
        assert lt in (0, 1) and rt in (0, 1)
        seen.add((lt, rt))
    assert len(seen) == 4
    l0, r0, _, _ = oracle.prg_expand(np.zeros(4, dtype=np.uint32))
    assert l0.any() and r0.any()


# ------------------------------------------------------------- key wire ----

def test_wire_length_rule_and_roundtrip():
    for d in (1, 8, 10, 20, 27, 31, 32):
        k0, k1 = oracle.keygen(d, (1 << d) - 1, seed=d)
        b0, b1 = oracle.key_to_bytes(k0), oracle.key_to_bytes(k1)
        assert len(b0) == len(b1) == 19 + 18 * d + 16
        assert b0[16] == 0 and b0[17] == 0 and b0[18] == d
        assert b1[16] == 1 and b1[17] == 1 and b1[18] == d
        assert oracle.key_to_bytes(oracle.key_from_bytes(b0)) == b0
        r = oracle.key_from_bytes(b1)
        for i in (0, 1, (1 << d) - 1):
            assert np.array_equal(oracle.eval_point(k1, i), oracle.eval_point(r, i))
    # Wire sizes at the benchmark domains: 395 B at d = 20, 521 B at d = 27, 593 B at d = 31.
    assert [oracle.key_wire_len(d) for d in (20, 27, 31)] == [395, 521, 593]


def test_wire_field_offsets_on_handbuilt_bytes():
    d = 3
    b = bytearray(19 + 18 * d + 16)
    b[0:16] = bytes(range(0x10, 0x20))
    b[16], b[17], b[18] = 0x05, 0x01, d
    for lvl in range(d):
        off = 19 + 18 * lvl
        b[off:off + 16] = bytes([0x40 + lvl] * 16)
        b[off + 16], b[off + 17] = 0xA0 + lvl, 0xB0 + lvl
    b[19 + 18 * d:] = bytes(range(0xE0, 0xF0))
    k = oracle.key_from_bytes(bytes(b))
    assert list(k.root_seed) == list(words_le(bytes(range(0x10, 0x20))))
    assert (k.root_t, k.party, k.domain_bits) == (5, 1, 3)
    for lvl in range(d):
        assert list(k.cw_seed[lvl]) == [0x40404040 + 0x01010101 * lvl] * 4
        assert (k.cw_t_left[lvl], k.cw_t_right[lvl]) == (0xA0 + lvl, 0xB0 + lvl)
    assert list(k.final_cw) == list(words_le(bytes(range(0xE0, 0xF0))))
    assert oracle.key_to_bytes(k) == bytes(b)


def test_wire_rejects_bad_lengths_and_domains():
    with pytest.raises(oracle.OracleError) as e:
        oracle.key_from_bytes(bytes(34))
    assert e.value.code == oracle.ERR_KEY_FORMAT
    good = oracle.keygen_bytes(8, 3, seed=1)[0]
    for bad in (good[:-1], good + b"\0"):
        with pytest.raises(oracle.OracleError) as e:
            oracle.key_from_bytes(bad)
        assert e.value.code == oracle.ERR_KEY_FORMAT
    b = bytearray(19 + 18 * 33 + 16)
    b[18] = 33
    with pytest.raises(oracle.OracleError) as e:
        oracle.key_from_bytes(bytes(b))
    assert e.value.code == oracle.ERR_DOMAIN
    # d = 0 is a valid wire key (one leaf): root seed, xor final_cw iff t == 1.
    z = bytearray(35)
    z[0:16] = bytes([7] * 16)
    z[16] = 1
    z[19:35] = bytes([0xFF] * 16)
    k = oracle.key_from_bytes(bytes(z))
    assert np.array_equal(oracle.eval_point(k, 0), np.full(4, 0xF8F8F8F8, dtype=np.uint32))


def test_keygen_rejects_out_of_range():
    for d, target in ((8, 256), (8, 300), (1, 2)):
        with pytest.raises(oracle.OracleError) as e:
            oracle.keygen(d, target, seed=0)
        assert e.value.code == oracle.ERR_INDEX
    for d in (0, 33):
        with pytest.raises(oracle.OracleError) as e:
            oracle.keygen(d, 0, seed=0)
        assert e.value.code == oracle.ERR_DOMAIN


# ------------------------------------------------------------ DPF eval -----

@pytest.mark.parametrize("d,target", [(1, 0), (1, 1), (2, 2), (8, 42), (10, 500),
                                      (12, 4095), (16, 32000)])
def test_two_party_xor_is_point_function_full_domain(d, target):
    k0, k1 = oracle.keygen(d, target, seed=1000 + d)
    x = oracle.full_eval(k0) ^ oracle.full_eval(k1)
    expect = np.zeros_like(x)
    expect[target] = 0xFFFFFFFF
    assert np.array_equal(x, expect)


def test_full_eval_equals_point_eval():
    # Pins the MSB-first path order: full_eval puts the left child first.
    for d, target in ((8, 100), (11, 1234)):
        k0, _ = oracle.keygen(d, target, seed=d)
        full = oracle.full_eval(k0)
        idx = range(1 << d) if d <= 8 else list(range(0, 1 << d, 37)) + [(1 << d) - 1]
        for i in idx:
            assert np.array_equal(full[i], oracle.eval_point(k0, i)), i


def test_eval_subtree_is_slice_of_full_eval():
    d = 12
    k0, k1 = oracle.keygen(d, 2048, seed=99)
    for k in (k0, k1):
        full = oracle.full_eval(k)
        for size in (1, 2, 64, 2048, 4096):
            for start in sorted({0, size, (1 << d) - size, 2048 // size * size}):
                if start + size > (1 << d):
                    continue
                assert np.array_equal(oracle.eval_subtree(k, start, size),
                                      full[start:start + size])
    with pytest.raises(oracle.OracleError):
        oracle.eval_subtree(k0, 3, 4)
    with pytest.raises(oracle.OracleError):
        oracle.eval_subtree(k0, 0, 3)


def test_boundary_targets():
    # dpf.rs:554-584
    for bits in (4, 8, 12, 16):
        mx = 1 << bits
        for target in (0, mx // 2, mx - 1):
            k0, k1 = oracle.keygen(bits, target, seed=bits * 7 + target)
            x = oracle.eval_point(k0, target) ^ oracle.eval_point(k1, target)
            assert np.array_equal(x, ones_mask())
            o = (target + 1) % mx
            assert not (oracle.eval_point(k0, o) ^ oracle.eval_point(k1, o)).any()


# ---------------------------------------------------------------- scans ----

def stamped_rows(n, w):
    # kernel.rs:1390-1403 pattern: 2-byte index stamp then (i + j) & 0xFF.
    i = np.arange(n, dtype=np.int64)[:, None]
    j = np.arange(w, dtype=np.int64)[None, :]
    rows = ((i + j) & 0xFF).astype(np.uint8)
    rows[:, 0] = (np.arange(n) & 0xFF).astype(np.uint8)
    rows[:, 1] = ((np.arange(n) >> 8) & 0xFF).astype(np.uint8)
    return rows


@pytest.mark.parametrize("n,w,d", [(1, 16, 0), (1, 32, 1), (255, 32, 8), (256, 48, 8),
                                   (500, 1024, 10), (2048, 32, 11), (2049, 16, 12),
                                   (4096, 32, 12), (5000, 64, 13)])
def test_brute_equals_tiled_and_recovers_target(n, w, d):
    rows = oracle.fill_rows(seed=n * 31 + w, first_row=0, n_rows=n, row_bytes=w)
    targets = [0, n - 1, n // 2]
    pairs = [oracle.keygen_bytes(d, t, seed=t + 17) for t in targets] if d > 0 else []
    if d == 0:
        # d = 0 pair by hand: equal root seeds, t = 0 / 1, final_cw all-ones.
        z0 = bytearray(35)
        z0[19:35] = bytes([0xFF] * 16)
        z1 = bytearray(z0)
        z1[16], z1[17] = 1, 1
        pairs = [(bytes(z0), bytes(z1))]
        targets = [0]
    for threads in (1, 3):
        a0 = oracle.scan_tiled([p[0] for p in pairs], rows, n_threads=threads)
        a1 = oracle.scan_tiled([p[1] for p in pairs], rows, n_threads=threads)
        for q, (k0, k1) in enumerate(pairs):
            assert np.array_equal(a0[q], oracle.scan_brute(k0, rows))
            assert np.array_equal(a1[q], oracle.scan_brute(k1, rows))
            assert np.array_equal(a0[q] ^ a1[q], rows[targets[q]])


def test_reference_page_pattern_recovery():
    # kernel.rs:1464-1497 and 1539-1561 at 4096-byte rows (the reference page).
    rows = stamped_rows(4096, 4096)
    targets = [2048, 100, 3000]
    pairs = [oracle.keygen_bytes(12, t, seed=t) for t in targets]
    a0 = oracle.scan_tiled([p[0] for p in pairs], rows, n_threads=4)
    a1 = oracle.scan_tiled([p[1] for p in pairs], rows, n_threads=4)
    for q, t in enumerate(targets):
        assert np.array_equal(a0[q] ^ a1[q], rows[t])


def test_scan_closed_forms():
    d, n, w = 10, 700, 64
    k0, k1 = oracle.keygen_bytes(d, 5, seed=3)
    zero = np.zeros((n, w), dtype=np.uint8)
    assert not oracle.scan_tiled([k0], zero).any()
    ones = np.full((n, w), 0xFF, dtype=np.uint8)
    masks = oracle.full_eval(oracle.key_from_bytes(k0))[:n]
    fold = np.bitwise_xor.reduce(masks, axis=0).astype("<u4").tobytes()
    assert oracle.scan_tiled([k0], ones)[0].tobytes() == fold * (w // 16)
    # A single non-zero row at r yields row & mask(r).
    one = np.zeros((n, w), dtype=np.uint8)
    one[123] = np.arange(w, dtype=np.uint8) * 3 + 1
    m = np.tile(masks[123].astype("<u4").view(np.uint8), w // 16)
    assert np.array_equal(oracle.scan_tiled([k1], one)[0] ^ oracle.scan_tiled([k0], one)[0],
                          np.zeros(w, dtype=np.uint8))
    assert np.array_equal(oracle.scan_tiled([k0], one)[0], one[123] & m)


def test_scan_rejects_bad_shapes():
    k0, _ = oracle.keygen_bytes(8, 1, seed=0)
    with pytest.raises(oracle.OracleError) as e:
        oracle.scan_tiled([k0], np.zeros((257, 32), dtype=np.uint8))
    assert e.value.code == oracle.ERR_ROWS
    with pytest.raises(oracle.OracleError) as e:
        oracle.scan_tiled([k0], np.zeros((4, 24), dtype=np.uint8))
    assert e.value.code == oracle.ERR_ROWS


def test_fuzzed_control_bytes_brute_equals_tiled():
    # Wire keys may carry control bytes other than 0/1; only `t == 1` (a byte
    # compare) triggers the correction word. Both evaluators must agree.
    rng = np.random.default_rng(5)
    d, n, w = 9, 512, 32
    rows = oracle.fill_rows(seed=11, first_row=0, n_rows=n, row_bytes=w)
    for trial in range(6):
        b = bytearray(oracle.keygen_bytes(d, int(rng.integers(0, n)), seed=trial)[trial % 2])
        b[16] = int(rng.choice([0, 1, 3, 0xFF]))
        for lvl in range(d):
            off = 19 + 18 * lvl
            b[off + 16] = int(rng.choice([0, 1, 2, 0x81]))
            b[off + 17] = int(rng.choice([0, 1, 3, 0xFE]))
        key = bytes(b)
        assert np.array_equal(oracle.scan_tiled([key], rows)[0], oracle.scan_brute(key, rows))


def test_delta_linearity():
    # scan(M xor D) == scan(M) xor scan_delta(D), with repeated indices.
    d, n, w = 11, 1500, 32
    rows = oracle.fill_rows(seed=5, first_row=0, n_rows=n, row_bytes=w)
    rng = np.random.default_rng(8)
    idx = rng.integers(0, n, size=300).astype(np.uint64)
    idx[:5] = idx[5]
    diffs = rng.integers(0, 256, size=(300, w), dtype=np.uint8)
    merged = rows.copy()
    for i, r in zip(idx, diffs):
        merged[int(i)] ^= r
    keys = [oracle.keygen_bytes(d, t, seed=t)[p] for t, p in ((7, 0), (1499, 1), (800, 0))]
    lhs = oracle.scan_tiled(keys, merged)
    rhs = oracle.scan_tiled(keys, rows) ^ oracle.scan_delta(keys, idx, diffs)
    assert np.array_equal(lhs, rhs)


# ------------------------------------------------------------ seeded data --

def test_splitmix64_known_answer():
    # First outputs of splitmix64 seeded with 0 (Vigna's reference generator).
    assert oracle.splitmix64_word(0, 0) == 0xE220A8397B1DCDAF
    assert oracle.splitmix64_word(0, 1) == 0x6E789E6AA1B965F4
    rows = oracle.fill_rows(seed=0, first_row=0, n_rows=1, row_bytes=16)
    assert rows[0, :8].tobytes() == (0xE220A8397B1DCDAF).to_bytes(8, "little")


def test_fill_rows_is_random_access():
    full = oracle.fill_rows(seed=42, first_row=0, n_rows=100, row_bytes=48)
    part = oracle.fill_rows(seed=42, first_row=37, n_rows=20, row_bytes=48)
    assert np.array_equal(full[37:57], part)


def test_lcg_jump_ahead_matches_sequential():
    # chunked_matrix.rs:83-93: state = state * A + 1, byte = state >> 56.
    a = 6364136223846793005
    seq = []
    state = 12345
    for _ in range(3000):
        state = (state * a + 1) & ((1 << 64) - 1)
        seq.append(state >> 56)
    assert oracle.lcg_fill(12345, 0, 3000).tolist() == seq
    assert oracle.lcg_fill(12345, 1234, 100).tolist() == seq[1234:1334]
    assert oracle.lcg_state_at(7, 0) == 7
    assert oracle.lcg_state_at(0, 1) == 1


# ------------------------------------------------------ fast oracle forms --
# oracle_fast.c restates the same scan with 16-lane vector ChaCha8 and wide
# folds; each form is pinned against the scalar functions above.

def fuzzed_key(d, target, seed, rng):
    """A keygen key whose control bytes are replaced by arbitrary bytes."""
    b = bytearray(oracle.keygen_bytes(d, target, seed=seed)[seed % 2])
    b[16] = int(rng.choice([0, 1, 2, 0xFF]))
    for lvl in range(d):
        off = 19 + 18 * lvl
        b[off + 16] = int(rng.choice([0, 1, 3, 0x81]))
        b[off + 17] = int(rng.choice([0, 1, 2, 0xFE]))
    return oracle.key_from_bytes(bytes(b))


@pytest.mark.parametrize("d", [1, 4, 5, 12, 17, 20])
def test_eval_subtree_fast_equals_scalar(d):
    rng = np.random.default_rng(d)
    keys = [oracle.keygen(d, int(rng.integers(0, 1 << d)), seed=s)[s % 2] for s in range(3)]
    keys.append(fuzzed_key(d, int(rng.integers(0, 1 << d)), 9, rng))
    for k in keys:
        for levels in sorted({0, 1, 3, 4, 5, min(d, 12), d}):
            if levels > d:
                continue
            size = 1 << levels
            for start in {0, ((1 << d) - size), int(rng.integers(0, 1 << (d - levels))) * size}:
                assert np.array_equal(oracle.eval_subtree_fast(k, start, size),
                                      oracle.eval_subtree(k, start, size)), (d, levels, start)


def test_eval_subtree_fast_refusals():
    k = oracle.keygen(8, 3, seed=1)[0]
    for start, size in [(0, 0), (0, 3), (2, 4), (0, 512)]:
        with pytest.raises(oracle.OracleError) as e:
            oracle.eval_subtree_fast(k, start, size)
        assert e.value.code == oracle.ERR_SUBTREE


def test_path_node_is_the_walk_of_eval():
    d = 14
    k = oracle.keygen(d, 9000, seed=4)[1]
    seed, t = oracle.path_node(k, 9000, d)
    leaf = oracle.eval_point(k, 9000)
    fcw = np.array(list(k.final_cw), dtype=np.uint32)
    assert np.array_equal(leaf, seed ^ fcw if t == 1 else seed)


@pytest.mark.parametrize("n,w,d", [(1, 16, 0), (7, 16, 3), (255, 32, 8), (4096, 32, 12),
                                   (5000, 48, 13), (9001, 64, 14), (3000, 144, 12),
                                   (20000, 1024, 15), (1 << 16, 32, 16)])
def test_fast_scans_equal_tiled(n, w, d):
    rng = np.random.default_rng(n + w)
    rows = oracle.fill_rows(seed=n, first_row=0, n_rows=n, row_bytes=w)
    targets = [int(rng.integers(0, max(n, 1))) for _ in range(3)]
    if d:
        keys = [oracle.keygen(d, t, seed=i)[i % 2] for i, t in enumerate(targets)]
        keys.append(fuzzed_key(d, targets[0], 5, rng))
    else:
        # d = 0 keys by hand: root t = 1 and 2 with a final correction word.
        z = bytearray(35)
        z[:16] = bytes(range(16))
        z[19:35] = bytes([0xA5] * 16)
        keys = []
        for t in (1, 2):
            z[16] = t
            keys.append(bytes(z))
    ref = oracle.scan_tiled(keys, rows, 3)
    assert np.array_equal(oracle.scan_fast(keys, rows, 3), ref)
    assert np.array_equal(oracle.scan_rows8(keys, rows, 3), ref)
    assert np.array_equal(oracle.scan_seeded(keys, n, n, w, n_threads=3), ref)
    # One thread and many threads agree (the per-thread folds XOR together).
    assert np.array_equal(oracle.scan_fast(keys, rows, 1), ref)
    if d:
        # Two-party reconstruction through the fast form alone.
        k0, k1 = oracle.keygen(d, targets[1], seed=77)
        a = oracle.scan_seeded([k0, k1], n, n, w, n_threads=2)
        assert np.array_equal(a[0] ^ a[1], rows[targets[1]])


def test_scan_seeded_row_ranges():
    # Answers over [a, b) equal the tiled scan of the matrix with every row
    # outside [a, b) zeroed, and the ranges XOR to the whole (ragged tile edges).
    n, w, d = 30000, 32, 15
    rows = oracle.fill_rows(seed=3, first_row=0, n_rows=n, row_bytes=w)
    keys = [oracle.keygen(d, t, seed=t)[t % 2] for t in (5, 17000, 29999)]
    cuts = [0, 1, 4095, 4096, 12289, 29999, 30000]
    parts = []
    for a, b in zip(cuts, cuts[1:]):
        got = oracle.scan_seeded(keys, 3, n, w, row_begin=a, row_end=b, n_threads=2)
        masked = np.zeros_like(rows)
        masked[a:b] = rows[a:b]
        assert np.array_equal(got, oracle.scan_tiled(keys, masked, 2)), (a, b)
        parts.append(got)
    whole = np.bitwise_xor.reduce(np.stack(parts), axis=0)
    assert np.array_equal(whole, oracle.scan_tiled(keys, rows, 2))


def test_fast_scan_refusals():
    k10 = oracle.keygen(10, 1, seed=1)[0]
    k11 = oracle.keygen(11, 1, seed=1)[0]
    rows = np.zeros((16, 32), dtype=np.uint8)
    with pytest.raises(oracle.OracleError) as e:
        oracle.scan_fast([k10, k11], rows)
    assert e.value.code == oracle.ERR_DOMAIN
    with pytest.raises(oracle.OracleError) as e:
        oracle.scan_seeded([k10], 1, 2048, 32)  # more rows than 2^d
    assert e.value.code == oracle.ERR_ROWS
    with pytest.raises(oracle.OracleError) as e:
        oracle.scan_seeded([k10], 1, 1024, 32, row_begin=10, row_end=5)
    assert e.value.code == oracle.ERR_ROWS
