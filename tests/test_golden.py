"""The oracle still reproduces the committed scan fixtures (regression pin)."""
import json
import os

import numpy as np

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "scan_fixtures.json")


def load():
    with open(GOLDEN) as f:
        return json.load(f)


def test_scan_fixtures_reproduce():
    fx = load()
    for c in fx["scan_cases"]:
        rows = oracle.fill_rows(c["data_seed"], 0, c["n_rows"], c["row_bytes"])
        keys = [bytes.fromhex(k) for k in c["keys"]]
        ans = oracle.scan_tiled(keys, rows, 2)
        assert [a.tobytes().hex() for a in ans] == c["answers"]
        for i, t in enumerate(c["targets"]):
            assert np.array_equal(ans[2 * i] ^ ans[2 * i + 1], rows[t])


def test_delta_fixtures_reproduce():
    for c in load()["delta_cases"]:
        keys = [bytes.fromhex(k) for k in c["keys"]]
        diffs = np.frombuffer(bytes.fromhex(c["diffs"]), np.uint8).reshape(-1, c["row_bytes"])
        ans = oracle.scan_delta(keys, np.array(c["row_idx"], dtype=np.uint64), diffs)
        assert [a.tobytes().hex() for a in ans] == c["answers"]
