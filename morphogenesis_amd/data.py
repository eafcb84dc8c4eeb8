"""Host-side regeneration of the device's seeded synthetic rows.

mg_matrix_fill_seeded writes 64-bit little-endian word j of the row-major
matrix as output j of a splitmix64 generator started at `seed`; this module
computes the same words with numpy, so any row can be recomputed on the host
without reading the device.
"""
import numpy as np

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
