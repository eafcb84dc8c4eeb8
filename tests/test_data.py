"""CPU tests of the row-data helpers: the 32-byte Compact account schema
(docs/DATA_STRUCTURES.md:17-26) as the reference client decodes it
(crates/morphogen-client/src/lib.rs:101-130; its tests at :528-562 give the
vectors below), and the host regeneration of the device's seeded rows.
"""
import numpy as np
import pytest

import oracle
from morphogenesis_amd.data import (AccountData, compact_row, compact_rows, seeded_rows)


def test_account_data_parses_compact_row():
    row = bytearray(32)
    row[15] = 1                              # balance = 1
    row[23] = 2                              # nonce = 2
    row[24:28] = bytes([0xDE, 0xAD, 0xBE, 0xEF])  # code id
    data = AccountData.from_bytes(bytes(row))
    assert (data.balance, data.nonce, data.code_id, data.code_hash) == (1, 2, 0xDEADBEEF, None)
    assert compact_row(1, 2, 0xDEADBEEF) == bytes(row)


def test_account_data_parses_full_row():
    row = bytearray(64)
    row[15] = 100
    row[23] = 5
    row[24:56] = b"\xaa" * 32
    data = AccountData.from_bytes(bytes(row))
    assert (data.balance, data.nonce, data.code_hash, data.code_id) == (100, 5, b"\xaa" * 32, None)


def test_other_lengths_do_not_parse():
    for n in (0, 31, 33, 63, 128, 4095):
        assert AccountData.from_bytes(bytes(n)) is None
    # A 4096-byte page is read as one Compact row at its start (lib.rs:104-105).
    page = compact_row(2**100 + 3, 9, 4) + bytes(4096 - 32)
    assert AccountData.from_bytes(page).balance == 2**100 + 3


def test_vectorised_rows_equal_single_rows():
    rng = np.random.default_rng(0)
    bal = rng.integers(0, 2**63, size=50, dtype=np.uint64)
    non = rng.integers(0, 2**63, size=50, dtype=np.uint64)
    cid = rng.integers(0, 2**32, size=50, dtype=np.uint64).astype(np.uint32)
    rows = compact_rows(bal, non, cid)
    for i in range(50):
        assert rows[i].tobytes() == compact_row(int(bal[i]), int(non[i]), int(cid[i]))
        a = AccountData.from_bytes(rows[i].tobytes())
        assert (a.balance, a.nonce, a.code_id) == (int(bal[i]), int(non[i]), int(cid[i]))
    with pytest.raises(ValueError):
        compact_row(1, 1, 1, padding=b"x")


def test_seeded_rows_equal_the_oracle_fill():
    for seed, first, n, w in ((5, 0, 10, 32), (7, 1000, 3, 1024), (0, 12345, 1, 16)):
        assert np.array_equal(seeded_rows(seed, first, n, w), oracle.fill_rows(seed, first, n, w))
