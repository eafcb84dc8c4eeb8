"""The autotuner's table and candidate grid (host logic only; no device)."""
import json

import pytest


@pytest.fixture(scope="module")
def at():
    from morphogenesis_amd import autotune
    return autotune


def test_keys_bucket_and_table_key(at):
    assert [at.keys_bucket(n) for n in (1, 2, 3, 4, 5, 24, 96, 384)] == [1, 2, 4, 4, 8, 32, 128,
                                                                          512]
    # log2 rows is the bit length of n - 1: 2^31 rows -> R31, 78,643,200 rows -> R27.
    assert at.table_key("gfx950", 32, 1 << 31, 24) == "gfx950|W32|R31|Q32"
    assert at.table_key("gfx950", 1024, 78_643_200, 3) == "gfx950|W1024|R27|Q4"


def test_lookup_reads_a_stored_setting(at, tmp_path):
    path = tmp_path / "t.json"
    assert at.lookup("gfx950", 1024, 78_643_200, 3, path=str(path)) is None
    entry = {"kernel": 5, "lane_depth": 6, "keys_per_wave": 3}
    path.write_text(json.dumps({at.table_key("gfx950", 1024, 78_643_200, 3): entry}))
    assert at.lookup("gfx950", 1024, 78_643_200, 3, path=str(path)) == entry
    # Same bucket (3 and 4 keys), other width or arch: no entry.
    assert at.lookup("gfx950", 1024, 78_643_200, 4, path=str(path)) == entry
    assert at.lookup("gfx950", 2048, 78_643_200, 3, path=str(path)) is None
    assert at.lookup("gfx942", 1024, 78_643_200, 3, path=str(path)) is None


def test_candidates_cover_each_kernel_for_its_widths(at):
    import morphogenesis_amd as mg
    narrow = at.candidates(32, 24)
    assert narrow[0] == {"kernel": mg.KERNEL_AUTO}
    assert any(c["kernel"] == mg.KERNEL_LANE_DFS for c in narrow)
    wide = at.candidates(1024, 3)
    assert wide[0] == {"kernel": mg.KERNEL_AUTO}
    assert not any(c["kernel"] == mg.KERNEL_LANE_DFS for c in wide)
    assert any(c["kernel"] == mg.KERNEL_STREAM_SPLIT and c["keys_per_wave"] == 3 for c in wide)
    # Never more keys per wave than keys, nor than the width allows.
    assert all(c.get("keys_per_wave", 1) <= 3 for c in wide)
    assert all(c.get("keys_per_wave", 1) <= (
        (4 if c.get("row_halves") else 3) if c["kernel"] == mg.KERNEL_STREAM_SPLIT else 2)
        for c in at.candidates(4096, 8))
    # Column halves only where rows are wider than 1 KiB; both ways there.
    assert not any("row_halves" in c for c in wide)
    halves = {c.get("row_halves") for c in at.candidates(2048, 4)
              if c["kernel"] == mg.KERNEL_STREAM_SPLIT}
    assert halves == {False, True}
    assert any(c.get("mask_depth") == 2 for c in wide)
    assert {c.get("late_loads") for c in narrow if c["kernel"] == mg.KERNEL_LANE_DFS} == {
        None, False, True}
    assert any(c.get("keys_per_wave") == 3 for c in at.candidates(4096, 8))
    assert any(c.get("keys_per_wave") == 4 for c in at.candidates(2048, 8))
    # Every stored field is one make_opts accepts.
    for c in narrow + wide + at.candidates(4096, 8):
        assert set(c) <= set(at.SETTING_KEYS)
        mg.make_opts(**c)
