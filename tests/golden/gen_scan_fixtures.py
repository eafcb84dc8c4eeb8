"""Writes scan_fixtures.json: seeded scan cases with oracle answers.

Self-generated by this project's C oracle (oracle/), which is pinned by
OpenSSL ChaCha20 and the DPF invariants in tests/test_oracle.py; no reference
code was run. Each case stores the data seed (rows = oracle.fill_rows), the
wire keys and the expected per-key answers, so the HIP path can be checked
against fixed bytes and the oracle against regressions.

Usage: python tests/golden/gen_scan_fixtures.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402

CASES = [  # n_rows, row_bytes, domain_bits, data seed
    (1, 32, 1, 1), (255, 32, 8, 2), (2048, 16, 11, 3), (2049, 32, 12, 4),
    (4096, 1024, 12, 5), (300, 4096, 9, 6), (5000, 48, 20, 7),
]


def main():
    out = {"note": "self-generated by the oracle; pinned by OpenSSL ChaCha20 + DPF invariants",
           "scan_cases": [], "delta_cases": []}
    for n, w, d, seed in CASES:
        rows = oracle.fill_rows(seed, 0, n, w)
        targets = sorted({0, n - 1, n // 2})
        pairs = [oracle.keygen_bytes(d, t, seed=seed * 1000 + t) for t in targets]
        keys = [k for p in pairs for k in p]
        ans = oracle.scan_tiled(keys, rows, 4)
        for i, t in enumerate(targets):
            assert np.array_equal(ans[2 * i] ^ ans[2 * i + 1], rows[t])
        out["scan_cases"].append({
            "n_rows": n, "row_bytes": w, "domain_bits": d, "data_seed": seed,
            "targets": targets, "keys": [k.hex() for k in keys],
            "answers": [a.tobytes().hex() for a in ans]})
    rng = np.random.default_rng(99)
    n, w, d = 4096, 32, 12
    idx = rng.integers(0, n, size=64).astype(np.uint64)
    idx[1] = idx[0]
    diffs = rng.integers(0, 256, size=(64, w), dtype=np.uint8)
    keys = list(oracle.keygen_bytes(d, int(idx[0]), seed=5))
    ans = oracle.scan_delta(keys, idx, diffs)
    out["delta_cases"].append({
        "n_rows": n, "row_bytes": w, "domain_bits": d, "data_seed": 8,
        "row_idx": [int(i) for i in idx], "diffs": diffs.tobytes().hex(),
        "keys": [k.hex() for k in keys], "answers": [a.tobytes().hex() for a in ans],
        "merged_answers": [a.tobytes().hex() for a in oracle.scan_tiled(
            keys, merge(oracle.fill_rows(8, 0, n, w), idx, diffs), 4)]})
    with open(os.path.join(HERE, "scan_fixtures.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


def merge(rows, idx, diffs):
    rows = rows.copy()
    for i, dlt in zip(idx, diffs):
        rows[int(i)] ^= dlt
    return rows


if __name__ == "__main__":
    main()
