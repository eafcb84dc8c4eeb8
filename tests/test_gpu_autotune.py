"""The autotuner on the device: a small grid tuned end to end, the winner
stored and read back, and the stored setting's answers equal to the oracle's."""
import json
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def test_tune_store_lookup_scan(tmp_path):
    import morphogenesis_amd as mg
    from morphogenesis_amd import autotune as at
    n, w, d = 9000, 1024, 14
    print(f"case autotune n={n} w={w} d={d}", flush=True)
    rows = oracle.fill_rows(seed=77, first_row=0, n_rows=n, row_bytes=w)
    s = mg.GpuScanner(0)
    try:
        m = mg.GpuPageMatrix.new(s, rows)
        pairs = [oracle.keygen_bytes(d, t, seed=t) for t in (0, 4321, n - 1)]
        keys = [p[0] for p in pairs]
        ref = oracle.scan_tiled(keys, rows, THREADS)
        grid = [dict(kernel=mg.KERNEL_AUTO),
                dict(kernel=mg.KERNEL_STREAM, lane_depth=5, keys_per_wave=2),
                dict(kernel=mg.KERNEL_STREAM_SPLIT, lane_depth=6, keys_per_wave=3),
                # d < kl + 6: the split kernel refuses it and the tuner skips it.
                dict(kernel=mg.KERNEL_STREAM_SPLIT, lane_depth=9, keys_per_wave=3)]
        logs = []
        best, table_rows = at.tune(s, m, keys, reps=2, grid=grid, log=logs.append)
        assert len(table_rows) == 3 and all(r["exact"] for r in table_rows)
        assert any(line.startswith("skip") for line in logs)
        assert best in table_rows and best["ms"] == min(r["ms"] for r in table_rows)

        path = str(tmp_path / "tuned.json")
        entry = {k: best[k] for k in at.SETTING_KEYS if k in best}
        with open(path, "w") as f:
            json.dump({at.table_key(at.device_arch(s), w, n, len(keys)): entry}, f)
        opts = at.tuned_opts(s, m, len(keys), path=path)
        assert opts is not None and opts.kernel == entry["kernel"]
        assert np.array_equal(s.scan_keys(m, keys, opts), ref)
        assert at.tuned_opts(s, m, 64, path=path) is None
        # Partner keys through the same setting reconstruct the rows.
        ans1 = s.scan_keys(m, [p[1] for p in pairs], opts)
        for q, t in enumerate((0, 4321, n - 1)):
            assert np.array_equal(ref[q] ^ ans1[q], rows[t])
        m.close()
    finally:
        s.close()
