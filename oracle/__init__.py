"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, by the smoke check and by the CPU
baseline leg of bench.py, always as the reference answers to compare against or the timed CPU baseline,
never as part of the product path. See oracle.h for the reference functions
each entry point restates and how each one is pinned.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
MAX_DOMAIN_BITS = 32

OK, ERR_KEY_FORMAT, ERR_DOMAIN, ERR_INDEX, ERR_SUBTREE, ERR_ROWS = range(6)


class OrcKey(ctypes.Structure):
    """In-memory ChaChaKey (dpf.rs:91-99), layout of orc_key in oracle.h."""
    _fields_ = [
        ("root_seed", ctypes.c_uint32 * 4),
        ("root_t", ctypes.c_uint8),
        ("party", ctypes.c_uint8),
        ("domain_bits", ctypes.c_uint8),
        ("cw_seed", (ctypes.c_uint32 * 4) * MAX_DOMAIN_BITS),
        ("cw_t_left", ctypes.c_uint8 * MAX_DOMAIN_BITS),
        ("cw_t_right", ctypes.c_uint8 * MAX_DOMAIN_BITS),
        ("final_cw", ctypes.c_uint32 * 4),
    ]


def build():
    """Compile liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        kp = ctypes.POINTER(OrcKey)
        sig = {
            "orc_chacha_block": (None, [u32p, ctypes.c_uint32, u32p, ctypes.c_int, u32p]),
            "orc_prg_expand": (None, [u32p, ctypes.c_int, u32p, u32p, u8p, u8p]),
            "orc_key_wire_len": (ctypes.c_size_t, [ctypes.c_int]),
            "orc_key_from_bytes": (ctypes.c_int, [u8p, ctypes.c_size_t, kp]),
            "orc_key_to_bytes": (ctypes.c_size_t, [kp, u8p]),
            "orc_eval": (None, [kp, ctypes.c_uint64, u32p]),
            "orc_eval_subtree": (ctypes.c_int, [kp, ctypes.c_uint64, ctypes.c_uint64, u32p]),
            "orc_full_eval": (ctypes.c_int, [kp, u32p]),
            "orc_keygen": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, kp, kp]),
            "orc_scan_brute": (ctypes.c_int, [kp, u8p, ctypes.c_uint64, ctypes.c_uint32, u8p]),
            "orc_scan_tiled": (ctypes.c_int, [kp, ctypes.c_int, u8p, ctypes.c_uint64,
                                              ctypes.c_uint32, ctypes.c_int, u8p]),
            "orc_scan_delta": (ctypes.c_int, [kp, ctypes.c_int, u64p, u8p, ctypes.c_uint64,
                                              ctypes.c_uint32, u8p]),
            "orc_path_node": (None, [kp, ctypes.c_uint64, ctypes.c_int, u32p, u8p]),
            "orc_eval_subtree_fast": (ctypes.c_int, [kp, ctypes.c_uint64, ctypes.c_uint64,
                                                     u32p]),
            "orc_scan_fast": (ctypes.c_int, [kp, ctypes.c_int, u8p, ctypes.c_uint64,
                                             ctypes.c_uint32, ctypes.c_int, u8p]),
            "orc_scan_rows8": (ctypes.c_int, [kp, ctypes.c_int, u8p, ctypes.c_uint64,
                                              ctypes.c_uint32, ctypes.c_int, u8p]),
            "orc_scan_seeded": (ctypes.c_int, [kp, ctypes.c_int, ctypes.c_uint64,
                                               ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                               ctypes.c_uint32, ctypes.c_int, u8p]),
            "orc_simd_isa": (ctypes.c_char_p, []),
            "orc_splitmix64_word": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint64]),
            "orc_fill_rows": (None, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                     ctypes.c_uint32, u8p]),
            "orc_lcg_fill": (None, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, u8p]),
            "orc_lcg_state_at": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint64]),
            "orc_siphash": (ctypes.c_uint64, [ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                              ctypes.c_uint64, u8p, ctypes.c_size_t]),
            "orc_cuckoo_indices": (ctypes.c_int, [u8p, ctypes.c_size_t, ctypes.c_uint64, u64p,
                                                  u64p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a, ctype=ctypes.c_uint8):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class OracleError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: oracle error {code}")
        self.code = code


def _check(rc, what):
    if rc != OK:
        raise OracleError(rc, what)


def chacha_block(key_words, counter, nonce_words=(0, 0, 0), rounds=8):
    k = np.asarray(key_words, dtype=np.uint32)
    n = np.asarray(nonce_words, dtype=np.uint32)
    out = np.zeros(16, dtype=np.uint32)
    lib().orc_chacha_block(_ptr(k, ctypes.c_uint32), counter, _ptr(n, ctypes.c_uint32),
                           rounds, _ptr(out, ctypes.c_uint32))
    return out


def prg_expand(seed_words, rounds=8):
    s = np.asarray(seed_words, dtype=np.uint32)
    left = np.zeros(4, dtype=np.uint32)
    right = np.zeros(4, dtype=np.uint32)
    lt = ctypes.c_uint8()
    rt = ctypes.c_uint8()
    lib().orc_prg_expand(_ptr(s, ctypes.c_uint32), rounds, _ptr(left, ctypes.c_uint32),
                         _ptr(right, ctypes.c_uint32), ctypes.byref(lt), ctypes.byref(rt))
    return left, right, lt.value, rt.value


def key_wire_len(domain_bits):
    return 19 + 18 * domain_bits + 16


def key_from_bytes(data: bytes) -> OrcKey:
    buf = np.frombuffer(bytes(data), dtype=np.uint8).copy()
    k = OrcKey()
    _check(lib().orc_key_from_bytes(_ptr(buf), len(buf), ctypes.byref(k)), "key_from_bytes")
    return k


def key_to_bytes(key: OrcKey) -> bytes:
    out = np.zeros(key_wire_len(key.domain_bits), dtype=np.uint8)
    n = lib().orc_key_to_bytes(ctypes.byref(key), _ptr(out))
    return out[:n].tobytes()


def keygen(domain_bits, target, seed):
    k0, k1 = OrcKey(), OrcKey()
    _check(lib().orc_keygen(domain_bits, target, seed, ctypes.byref(k0), ctypes.byref(k1)),
           "keygen")
    return k0, k1


def keygen_bytes(domain_bits, target, seed):
    k0, k1 = keygen(domain_bits, target, seed)
    return key_to_bytes(k0), key_to_bytes(k1)


def eval_point(key: OrcKey, index: int) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_eval(ctypes.byref(key), index, _ptr(out, ctypes.c_uint32))
    return out


def eval_subtree(key: OrcKey, start: int, size: int) -> np.ndarray:
    out = np.zeros((size, 4), dtype=np.uint32)
    _check(lib().orc_eval_subtree(ctypes.byref(key), start, size, _ptr(out, ctypes.c_uint32)),
           "eval_subtree")
    return out


def full_eval(key: OrcKey) -> np.ndarray:
    out = np.zeros((1 << key.domain_bits, 4), dtype=np.uint32)
    _check(lib().orc_full_eval(ctypes.byref(key), _ptr(out, ctypes.c_uint32)), "full_eval")
    return out


def _keys_array(keys):
    arr = (OrcKey * len(keys))()
    for i, k in enumerate(keys):
        arr[i] = k if isinstance(k, OrcKey) else key_from_bytes(k)
    return arr


def scan_brute(key, rows: np.ndarray) -> np.ndarray:
    k = key if isinstance(key, OrcKey) else key_from_bytes(key)
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    n, w = rows.shape
    out = np.zeros(w, dtype=np.uint8)
    _check(lib().orc_scan_brute(ctypes.byref(k), _ptr(rows), n, w, _ptr(out)), "scan_brute")
    return out


def scan_tiled(keys, rows: np.ndarray, n_threads=1) -> np.ndarray:
    arr = _keys_array(keys)
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    n, w = rows.shape
    out = np.zeros((len(keys), w), dtype=np.uint8)
    _check(lib().orc_scan_tiled(arr, len(keys), _ptr(rows), n, w, n_threads, _ptr(out)),
           "scan_tiled")
    return out


def _scan_rows_fn(fn, name, keys, rows, n_threads):
    arr = _keys_array(keys)
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    n, w = rows.shape
    out = np.zeros((len(keys), w), dtype=np.uint8)
    _check(fn(arr, len(keys), _ptr(rows), n, w, n_threads, _ptr(out)), name)
    return out


def scan_fast(keys, rows: np.ndarray, n_threads=1) -> np.ndarray:
    """scan_tiled's answers from the 16-lane vector ChaCha8 path (keys share one domain)."""
    return _scan_rows_fn(lib().orc_scan_fast, "scan_fast", keys, rows, n_threads)


def scan_rows8(keys, rows: np.ndarray, n_threads=1) -> np.ndarray:
    """scan_tiled's answers in the reference server scan's eight-rows-per-step shape."""
    return _scan_rows_fn(lib().orc_scan_rows8, "scan_rows8", keys, rows, n_threads)


def scan_seeded(keys, seed, n_rows, row_bytes, row_begin=0, row_end=None, n_threads=1):
    """Answers over rows [row_begin, row_end) of the n_rows x row_bytes matrix that
    fill_rows(seed, ...) generates, regenerated tile by tile (no host copy)."""
    arr = _keys_array(keys)
    row_end = n_rows if row_end is None else row_end
    out = np.zeros((len(keys), row_bytes), dtype=np.uint8)
    _check(lib().orc_scan_seeded(arr, len(keys), seed, n_rows, row_begin, row_end, row_bytes,
                                 n_threads, _ptr(out)), "scan_seeded")
    return out


def eval_subtree_fast(key: OrcKey, start: int, size: int) -> np.ndarray:
    out = np.zeros((size, 4), dtype=np.uint32)
    _check(lib().orc_eval_subtree_fast(ctypes.byref(key), start, size,
                                       _ptr(out, ctypes.c_uint32)), "eval_subtree_fast")
    return out


def path_node(key: OrcKey, index: int, levels: int):
    seed = np.zeros(4, dtype=np.uint32)
    t = ctypes.c_uint8()
    lib().orc_path_node(ctypes.byref(key), index, levels, _ptr(seed, ctypes.c_uint32),
                        ctypes.byref(t))
    return seed, t.value


def simd_isa() -> str:
    return lib().orc_simd_isa().decode()


def scan_delta(keys, row_idx, diffs: np.ndarray) -> np.ndarray:
    arr = _keys_array(keys)
    idx = np.ascontiguousarray(row_idx, dtype=np.uint64)
    diffs = np.ascontiguousarray(diffs, dtype=np.uint8)
    n, w = diffs.shape
    out = np.zeros((len(keys), w), dtype=np.uint8)
    _check(lib().orc_scan_delta(arr, len(keys), _ptr(idx, ctypes.c_uint64), _ptr(diffs), n, w,
                                _ptr(out)), "scan_delta")
    return out


def splitmix64_word(seed, j):
    return lib().orc_splitmix64_word(seed, j)


def fill_rows(seed, first_row, n_rows, row_bytes) -> np.ndarray:
    out = np.zeros((n_rows, row_bytes), dtype=np.uint8)
    lib().orc_fill_rows(seed, first_row, n_rows, row_bytes, _ptr(out))
    return out


def lcg_fill(seed, first_byte, n_bytes) -> np.ndarray:
    out = np.zeros(n_bytes, dtype=np.uint8)
    lib().orc_lcg_fill(seed, first_byte, n_bytes, _ptr(out))
    return out


def lcg_state_at(seed, steps):
    return lib().orc_lcg_state_at(seed, steps)


def siphash(msg: bytes, c_rounds=1, d_rounds=3, k0=0, k1=0) -> int:
    buf = np.frombuffer(bytes(msg) or b"\0", dtype=np.uint8).copy()
    return int(lib().orc_siphash(c_rounds, d_rounds, k0, k1, _ptr(buf), len(msg)))


def cuckoo_indices(key: bytes, num_rows: int, seeds=(0x12345678, 0x9ABCDEF0, 0xFEDCBA98)):
    """CuckooAddresser::with_seeds(num_rows, seeds).hash_indices(key) (cuckoo.rs:25-40);
    the default seeds are CuckooAddresser::new's."""
    buf = np.frombuffer(bytes(key) or b"\0", dtype=np.uint8).copy()
    sd = np.array(seeds, dtype=np.uint64)
    out = np.zeros(3, dtype=np.uint64)
    _check(lib().orc_cuckoo_indices(_ptr(buf), len(key), num_rows, _ptr(sd, ctypes.c_uint64),
                                    _ptr(out, ctypes.c_uint64)), "cuckoo_indices")
    return [int(x) for x in out]

