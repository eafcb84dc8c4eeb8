"""Host-side row data: the device's seeded synthetic rows, and the
reference's 32-byte Compact account rows.

mg_matrix_fill_seeded writes 64-bit little-endian word j of the row-major
matrix as output j of a splitmix64 generator started at `seed`; this module
computes the same words with numpy, so any row can be recomputed on the host
without reading the device.

Compact rows (docs/DATA_STRUCTURES.md:17-26) are Balance u128 | Nonce u64 |
CodeID u32 | 4 padding bytes, big-endian as the client decodes them
(crates/morphogen-client/src/lib.rs:101-130 AccountData::from_bytes).
"""
from dataclasses import dataclass
from typing import Optional

import numpy as np

COMPACT_ROW_BYTES = 32

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64_words(seed: int, first: int, n: int) -> np.ndarray:
    j = np.arange(first + 1, first + 1 + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + j * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def seeded_rows(seed: int, first_row: int, n_rows: int, row_bytes: int) -> np.ndarray:
    words = row_bytes // 8
    w = splitmix64_words(seed, first_row * words, n_rows * words)
    return w.astype("<u8").view(np.uint8).reshape(n_rows, row_bytes)


def compact_row(balance: int, nonce: int, code_id: int = 0, padding: bytes = bytes(4)) -> bytes:
    """One 32-byte Compact row (docs/DATA_STRUCTURES.md:17-26)."""
    if len(padding) != 4:
        raise ValueError("padding is 4 bytes")
    return (balance.to_bytes(16, "big") + nonce.to_bytes(8, "big") +
            code_id.to_bytes(4, "big") + bytes(padding))


def compact_rows(balances, nonces, code_ids) -> np.ndarray:
    """n Compact rows at once, as an (n, 32) uint8 array (balances below 2^64)."""
    b = np.asarray(balances, dtype=np.uint64)
    n = b.shape[0]
    out = np.zeros((n, COMPACT_ROW_BYTES), dtype=np.uint8)
    out[:, 8:16] = b.astype(">u8").view(np.uint8).reshape(n, 8)
    out[:, 16:24] = np.asarray(nonces, dtype=np.uint64).astype(">u8").view(np.uint8).reshape(n, 8)
    out[:, 24:28] = np.asarray(code_ids, dtype=np.uint32).astype(">u4").view(np.uint8).reshape(n, 4)
    return out


@dataclass
class AccountData:
    """AccountData (crates/morphogen-client/src/lib.rs:86-92)."""
    nonce: int
    balance: int
    code_hash: Optional[bytes]
    code_id: Optional[int]

    @staticmethod
    def from_bytes(row: bytes) -> Optional["AccountData"]:
        """AccountData::from_bytes (lib.rs:101-130): a 32-byte Compact row (or a
        4096-byte page read as one), or a 64-byte Full row; else None."""
        row = bytes(row)
        if len(row) in (32, 4096):
            return AccountData(int.from_bytes(row[16:24], "big"), int.from_bytes(row[0:16], "big"),
                               None, int.from_bytes(row[24:28], "big"))
        if len(row) == 64:
            return AccountData(int.from_bytes(row[16:24], "big"), int.from_bytes(row[0:16], "big"),
                               row[24:56], None)
        return None

