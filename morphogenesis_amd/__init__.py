"""MI355X-native fused ChaCha8-DPF masked-XOR PIR scan.

Python mirror of the reference's GPU scan interface
(crates/morphogen-gpu-dpf/src/kernel.rs GpuScanner, src/storage.rs
GpuPageMatrix, src/dpf.rs ChaChaKey wire format), driving the C ABI of
libmorphogen_hip.so (include/morphogen_hip.h) through ctypes. Every scan runs
on the GPU; there is no CPU fallback, and importing this package fails if the
library has not been built.
"""
import ctypes
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmorphogen_hip.so")

KEYS_PER_QUERY = 3  # cuckoo locations per query (kernel.rs KEYS_PER_QUERY)

MG_OK = 0
MG_ERR_INVALID_ARG = 1
MG_ERR_KEY_FORMAT = 2
MG_ERR_DOMAIN = 3
MG_ERR_ROWS = 4
MG_ERR_OUT_OF_MEMORY = 5
MG_ERR_DEVICE = 6
MG_ERR_NO_DEVICE = 7
MG_ERR_BUSY = 8
MG_ERR_STATE = 9

(KERNEL_AUTO, KERNEL_LANE_DFS, KERNEL_STREAM, KERNEL_POINT, KERNEL_SUFFIX,
 KERNEL_STREAM_SPLIT) = range(6)
KERNEL_NAMES = {KERNEL_LANE_DFS: "lane_dfs", KERNEL_STREAM: "stream", KERNEL_POINT: "point",
                KERNEL_SUFFIX: "suffix", KERNEL_STREAM_SPLIT: "stream_split"}


class MgError(RuntimeError):
    """Error returned through the C ABI (status code + mg_last_error text)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


class InvalidKeyFormat(MgError):
    """GpuDpfError::InvalidKeyFormat (dpf.rs:14)."""


class InvalidDomainBits(MgError):
    """GpuDpfError::InvalidDomainBits (dpf.rs:11)."""


class RowsError(MgError):
    """More rows than 2^d, bad row width or out-of-range rows."""


class NoDevice(MgError):
    pass


_ERRORS = {MG_ERR_KEY_FORMAT: InvalidKeyFormat, MG_ERR_DOMAIN: InvalidDomainBits,
           MG_ERR_ROWS: RowsError, MG_ERR_NO_DEVICE: NoDevice}


class KeyInfo(ctypes.Structure):
    _fields_ = [("domain_bits", ctypes.c_uint32), ("party", ctypes.c_uint32),
                ("root_t", ctypes.c_uint32), ("wire_len", ctypes.c_uint32)]


class Timing(ctypes.Structure):
    """KernelTiming (kernel.rs:1213-1228), per call, in milliseconds."""
    _fields_ = [("h2d_ms", ctypes.c_double), ("kernel_ms", ctypes.c_double),
                ("d2h_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("scan_kernel_ms", ctypes.c_double), ("kernel", ctypes.c_uint32),
                ("lane_depth", ctypes.c_uint32), ("keys_per_wave", ctypes.c_uint32),
                ("waves", ctypes.c_uint32), ("stack_global", ctypes.c_uint32),
                ("dfs_waves", ctypes.c_uint32)]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["kernel"] = KERNEL_NAMES.get(self.kernel, str(self.kernel))
        return d


class ScanOpts(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_uint32), ("lane_depth", ctypes.c_uint32),
                ("keys_per_wave", ctypes.c_uint32), ("waves_per_key", ctypes.c_uint32),
                ("waves_per_block", ctypes.c_uint32), ("use_graph", ctypes.c_uint32),
                ("row_begin", ctypes.c_uint64), ("row_end", ctypes.c_uint64),
                ("delta", ctypes.c_void_p), ("stack_global", ctypes.c_uint32),
                ("row_loads", ctypes.c_uint32), ("row_halves", ctypes.c_uint32),
                ("mask_depth", ctypes.c_uint32)]


class WireOpts(ctypes.Structure):
    _fields_ = [("max_body_bytes", ctypes.c_uint64), ("max_batch", ctypes.c_uint32),
                ("timeout_ms", ctypes.c_uint32)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("segments", ctypes.c_uint32), ("lane_depth", ctypes.c_uint32),
                ("max_active_keys", ctypes.c_uint32), ("domain_bits", ctypes.c_uint32),
                ("hbm_reserve_bytes", ctypes.c_uint64), ("max_delta_entries", ctypes.c_uint64),
                ("double_buffer", ctypes.c_uint32), ("deadline_ms", ctypes.c_uint32),
                ("use_graph", ctypes.c_uint32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("queue_ms", ctypes.c_double), ("service_ms", ctypes.c_double),
                ("latency_ms", ctypes.c_double), ("epoch", ctypes.c_uint64),
                ("n_keys", ctypes.c_uint32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("keys_completed", ctypes.c_uint64), ("sweep_steps", ctypes.c_uint64),
                ("epoch", ctypes.c_uint64), ("delta_entries", ctypes.c_uint64),
                ("rebases", ctypes.c_uint64), ("active_keys", ctypes.c_uint32),
                ("queued_requests", ctypes.c_uint32), ("max_active_keys", ctypes.c_uint32),
                ("segments", ctypes.c_uint32), ("busy_ms", ctypes.c_double),
                ("snapshots", ctypes.c_uint64), ("n_rows", ctypes.c_uint64),
                ("rejected", ctypes.c_uint64), ("abandoned", ctypes.c_uint64),
                ("key_step_ms", ctypes.c_double), ("step_fixed_ms", ctypes.c_double),
                ("graph_launches", ctypes.c_uint64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C {_HERE}` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    vp, u8p, u64p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_uint64)
    szp = ctypes.POINTER(ctypes.c_size_t)
    sigs = {
        "mg_abi_version": (ctypes.c_int, []),
        "mg_last_error": (ctypes.c_char_p, []),
        "mg_device_open": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(vp)]),
        "mg_device_close": (ctypes.c_int, [vp]),
        "mg_device_set_workspaces": (ctypes.c_int, [vp, ctypes.c_uint32]),
        "mg_device_info": (ctypes.c_int, [vp, u64p, u64p, ctypes.POINTER(ctypes.c_uint32),
                                          ctypes.c_char_p, ctypes.c_size_t]),
        "mg_matrix_create": (ctypes.c_int, [vp, ctypes.c_uint64, ctypes.c_uint32,
                                            ctypes.POINTER(vp)]),
        "mg_matrix_upload": (ctypes.c_int, [vp, ctypes.c_uint64, vp, ctypes.c_uint64]),
        "mg_matrix_download": (ctypes.c_int, [vp, ctypes.c_uint64, vp, ctypes.c_uint64]),
        "mg_matrix_fill_seeded": (ctypes.c_int, [vp, ctypes.c_uint64]),
        "mg_matrix_fill_pattern": (ctypes.c_int, [vp, ctypes.c_uint64]),
        "mg_matrix_info": (ctypes.c_int, [vp, u64p, ctypes.POINTER(ctypes.c_uint32)]),
        "mg_matrix_copy": (ctypes.c_int, [vp, vp]),
        "mg_matrix_destroy": (ctypes.c_int, [vp]),
        "mg_matrix_load_file": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_uint32,
                                               ctypes.c_uint32, ctypes.POINTER(vp)]),
        "mg_keygen": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, u8p,
                                     u8p]),
        "mg_key_parse": (ctypes.c_int, [u8p, ctypes.c_size_t, ctypes.POINTER(KeyInfo)]),
        "mg_key_eval": (ctypes.c_int, [vp, u8p, ctypes.c_size_t, u64p, ctypes.c_uint64, u8p]),
        "mg_key_eval_subtree": (ctypes.c_int, [vp, u8p, ctypes.c_size_t, ctypes.c_uint64,
                                               ctypes.c_uint64, u8p]),
        "mg_key_full_eval": (ctypes.c_int, [vp, u8p, ctypes.c_size_t, u8p, ctypes.c_uint64]),
        "mg_scan": (ctypes.c_int, [vp, vp, u8p, szp, ctypes.c_uint32, u8p,
                                   ctypes.POINTER(ScanOpts), ctypes.POINTER(Timing)]),
        "mg_scan_noalloc": (ctypes.c_int, [vp, vp, u8p, szp, ctypes.c_uint32, u8p,
                                           ctypes.POINTER(ScanOpts), ctypes.POINTER(Timing)]),
        "mg_scan_per_query": (ctypes.c_int, [vp, vp, u8p, szp, ctypes.c_uint32, ctypes.c_uint32,
                                             u8p, ctypes.POINTER(ScanOpts),
                                             ctypes.POINTER(Timing)]),
        "mg_delta_create": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.c_uint64,
                                           ctypes.POINTER(vp)]),
        "mg_delta_set": (ctypes.c_int, [vp, u64p, u8p, ctypes.c_uint64, ctypes.c_uint64]),
        "mg_delta_info": (ctypes.c_int, [vp, u64p, u64p]),
        "mg_delta_destroy": (ctypes.c_int, [vp]),
        "mg_scan_delta": (ctypes.c_int, [vp, vp, u8p, szp, ctypes.c_uint32, u8p]),
        "mg_matrix_apply_delta": (ctypes.c_int, [vp, vp]),
        "mg_server_create": (ctypes.c_int, [vp, vp, ctypes.POINTER(ServerOpts),
                                            ctypes.POINTER(vp)]),
        "mg_server_submit": (ctypes.c_int, [vp, u8p, szp, ctypes.c_uint32, u64p]),
        "mg_server_wait": (ctypes.c_int, [vp, ctypes.c_uint64, u8p, ctypes.c_uint32,
                                          ctypes.POINTER(RequestInfo)]),
        "mg_server_update": (ctypes.c_int, [vp, u64p, u8p, ctypes.c_uint64, ctypes.c_uint64]),
        "mg_server_stats_get": (ctypes.c_int, [vp, ctypes.POINTER(ServerStats)]),
        "mg_server_snapshot": (ctypes.c_int, [vp, vp, ctypes.c_uint64]),
        "mg_server_abandon": (ctypes.c_int, [vp, ctypes.c_uint64]),
        "mg_server_inject_failure": (ctypes.c_int, [vp, ctypes.c_int]),
        "mg_server_destroy": (ctypes.c_int, [vp]),
        "mg_cuckoo_hash_indices": (ctypes.c_int, [u8p, ctypes.c_size_t, ctypes.c_uint64, u64p,
                                                  u64p]),
        "mg_page_query": (ctypes.c_int, [u8p, ctypes.c_size_t, ctypes.c_uint64, u64p,
                                         ctypes.c_uint32, ctypes.c_uint64, u8p, u8p, u64p,
                                         ctypes.POINTER(ctypes.c_uint32)]),
        "mg_page_aggregate": (ctypes.c_int, [ctypes.c_uint64, u8p, ctypes.c_uint64, u8p,
                                             ctypes.c_size_t, ctypes.c_uint32, u8p]),
        "mg_page_extract_row": (ctypes.c_int, [u8p, ctypes.c_size_t, ctypes.c_uint32, u8p]),
        "mg_wire_response_capacity": (ctypes.c_size_t, [ctypes.c_uint32, ctypes.c_uint32]),
        "mg_wire_parse": (ctypes.c_int, [ctypes.POINTER(WireOpts), ctypes.c_char_p,
                                         ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                         ctypes.c_size_t, u8p, ctypes.c_size_t, szp,
                                         ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                         ctypes.POINTER(ctypes.c_int)]),
        "mg_wire_format": (ctypes.c_int, [ctypes.c_uint64, u8p, ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t, szp]),
        "mg_wire_handle": (ctypes.c_int, [vp, ctypes.POINTER(WireOpts), ctypes.c_char_p,
                                          ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                          ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t, szp,
                                          ctypes.POINTER(ctypes.c_int)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


lib = _load()


def _check(rc: int):
    if rc != MG_OK:
        msg = (lib.mg_last_error() or b"").decode(errors="replace")
        raise _ERRORS.get(rc, MgError)(rc, msg)


def _u8(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def _pack_keys(keys: Sequence[bytes]):
    blob = np.frombuffer(b"".join(bytes(k) for k in keys) or b"\0", dtype=np.uint8).copy()
    lens = (ctypes.c_size_t * max(1, len(keys)))(*[len(k) for k in keys])
    return blob, lens


def parse_key(data: bytes) -> KeyInfo:
    """ChaChaKey::from_bytes validation (dpf.rs:277-316) through the C ABI."""
    buf = np.frombuffer(bytes(data) or b"\0", dtype=np.uint8).copy()
    info = KeyInfo()
    _check(lib.mg_key_parse(_u8(buf), len(data), ctypes.byref(info)))
    return info


def generate_keys(domain_bits: int, target: int, seed: int):
    """generate_chacha_dpf_keys (dpf.rs:322-432) with seeded root seeds:
    returns the two parties' wire keys."""
    n = 19 + 18 * domain_bits + 16 if 0 < domain_bits <= 32 else 1
    k0 = np.zeros(n, dtype=np.uint8)
    k1 = np.zeros(n, dtype=np.uint8)
    rc = lib.mg_keygen(domain_bits, target, seed, _u8(k0), _u8(k1))
    if rc != MG_OK:
        raise _ERRORS.get(rc, MgError)(rc, f"keygen(d={domain_bits}, target={target})")
    return k0.tobytes(), k1.tobytes()


@dataclass
class PirResult:
    """PirResult (kernel.rs:1230-1240) generalised to W-byte rows: one answer
    per key of the query; verif fields stay zero as in the reference build
    (verifiable PIR compiled out)."""
    pages: List[bytes]
    verifs: List[bytes] = field(default_factory=list)
    timing: Optional[dict] = None

    @property
    def page0(self):
        return self.pages[0]

    @property
    def page1(self):
        return self.pages[1]

    @property
    def page2(self):
        return self.pages[2]


def make_opts(kernel=KERNEL_AUTO, lane_depth=0, keys_per_wave=0, waves_per_key=0,
              waves_per_block=0, use_graph=False, row_begin=0, row_end=0, delta=None,
              stack_global=None, late_loads=None, row_halves=None, mask_depth=0):
    """stack_global: lane-DFS stack levels kept in global memory (None = tuned);
    late_loads: lane-DFS rows loaded after a leaf pair's expansions (True) or
    before them (False); None = tuned. row_halves: split kernel on rows over
    1 KiB, two pairs per row, one per column half of 64 (rows up to 2 KiB) or
    128 chunks (True), or one pair per row (False); None = tuned (halves on
    rows over 2 KiB from 4 keys). mask_depth: split kernel, leaf-pair steps of masks a
    DFS wave may run ahead of its streaming wave (2-8; 0 = tuned)."""
    o = ScanOpts()
    o.kernel = kernel
    o.lane_depth = lane_depth
    o.keys_per_wave = keys_per_wave
    o.waves_per_key = waves_per_key
    o.waves_per_block = waves_per_block
    o.use_graph = 1 if use_graph else 0
    o.row_begin = row_begin
    o.row_end = row_end
    o.delta = delta._h if delta is not None else None
    o.stack_global = 0 if stack_global is None else 1 + stack_global
    o.row_loads = 0 if late_loads is None else (2 if late_loads else 1)
    o.row_halves = 0 if row_halves is None else (2 if row_halves else 1)
    o.mask_depth = mask_depth
    return o


class GpuScanner:
    """GpuScanner (kernel.rs:158-171): one gfx950 device and its pool of scan
    workspaces (a stream with its own buffers each); calls from several threads
    run concurrently on distinct ones."""

    def __init__(self, device_ord: int = 0):
        h = ctypes.c_void_p()
        _check(lib.mg_device_open(device_ord, ctypes.byref(h)))
        self._h = h
        self.device_ord = device_ord

    def close(self):
        """Releases the device; refused (MgError, MG_ERR_STATE) while matrices, deltas or
        request loops opened on it are still open."""
        if self._h:
            _check(lib.mg_device_close(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_workspaces(self, n: int):
        """Scan workspaces (streams) calls spread over, 1-8 (default 4)."""
        _check(lib.mg_device_set_workspaces(self._h, n))

    def info(self):
        fr, tot = ctypes.c_uint64(), ctypes.c_uint64()
        cus = ctypes.c_uint32()
        name = ctypes.create_string_buffer(256)
        _check(lib.mg_device_info(self._h, ctypes.byref(fr), ctypes.byref(tot), ctypes.byref(cus),
                                  name, 256))
        return {"free_bytes": fr.value, "total_bytes": tot.value, "compute_units": cus.value,
                "name": name.value.decode()}

    # -- flat key list (the C ABI's native form) --------------------------
    def scan_keys(self, db: "GpuPageMatrix", keys: Sequence[bytes], opts: ScanOpts = None,
                  with_timing=False):
        """Answers of n keys (n x W uint8) from one pass over the matrix."""
        n = len(keys)
        out = np.zeros((max(n, 1), db.row_bytes), dtype=np.uint8)
        blob, lens = _pack_keys(keys)
        t = Timing()
        _check(lib.mg_scan(self._h, db._h, _u8(blob), lens, n, _u8(out),
                           ctypes.byref(opts) if opts is not None else None, ctypes.byref(t)))
        out = out[:n]
        return (out, t.as_dict()) if with_timing else out

    def scan_keys_noalloc(self, db: "GpuPageMatrix", keys: Sequence[bytes],
                          opts: ScanOpts = None, with_timing=False):
        """scan_keys through mg_scan_noalloc: the same answers, with every per-call
        device buffer inside the library's fixed scratch array (MG_NOALLOC_SCRATCH_BYTES
        per GPU, at most MG_NOALLOC_MAX_KEYS keys)."""
        n = len(keys)
        out = np.zeros((max(n, 1), db.row_bytes), dtype=np.uint8)
        blob, lens = _pack_keys(keys)
        t = Timing()
        _check(lib.mg_scan_noalloc(self._h, db._h, _u8(blob), lens, n, _u8(out),
                                   ctypes.byref(opts) if opts is not None else None,
                                   ctypes.byref(t)))
        out = out[:n]
        return (out, t.as_dict()) if with_timing else out

    def scan_keys_per_query(self, db, keys, keys_per_query=KEYS_PER_QUERY, opts=None,
                            with_timing=False):
        """One matrix pass per query (the reference's multistream mode)."""
        n = len(keys)
        out = np.zeros((max(n, 1), db.row_bytes), dtype=np.uint8)
        blob, lens = _pack_keys(keys)
        t = Timing()
        _check(lib.mg_scan_per_query(self._h, db._h, _u8(blob), lens, n, keys_per_query,
                                     _u8(out), ctypes.byref(opts) if opts is not None else None,
                                     ctypes.byref(t)))
        out = out[:n]
        return (out, t.as_dict()) if with_timing else out

    # -- reference-shaped entry points -----------------------------------
    def scan(self, db, keys: Sequence[bytes], opts=None) -> PirResult:
        """GpuScanner::scan (kernel.rs:927-935): one query of 3 keys."""
        return self.scan_batch(db, [list(keys)], opts)[0]

    def scan_batch(self, db, queries: Sequence[Sequence[bytes]], opts=None) -> List[PirResult]:
        """GpuScanner::scan_batch (kernel.rs:685-700): all queries in one pass."""
        flat = [k for q in queries for k in q]
        if not flat:
            return []
        out, t = self.scan_keys(db, flat, opts, with_timing=True)
        res, i = [], 0
        for q in queries:
            pages = [out[i + j].tobytes() for j in range(len(q))]
            res.append(PirResult(pages, [bytes(16)] * len(q), t))
            i += len(q)
        return res

    def scan_batch_optimized_tiled(self, db, queries, tile_size: int, use_graph=False):
        """scan_batch_optimized_tiled(_with_graph) (kernel.rs:1124-1162): at
        most tile_size queries per pass."""
        tile = max(1, int(tile_size))
        out = []
        for i in range(0, len(queries), tile):
            out += self.scan_batch(db, queries[i:i + tile], make_opts(use_graph=use_graph))
        return out

    def scan_batch_single_query(self, db, queries, use_graph=False):
        """scan_batch_single_query_multistream_optimized(_with_graph)
        (kernel.rs:955-1097): one pass per query."""
        flat = [k for q in queries for k in q]
        if not flat:
            return []
        per = len(queries[0])
        out, t = self.scan_keys_per_query(db, flat, per, make_opts(use_graph=use_graph),
                                          with_timing=True)
        return [PirResult([out[i * per + j].tobytes() for j in range(per)], [bytes(16)] * per, t)
                for i in range(len(queries))]

    def eval(self, key: bytes, indices) -> np.ndarray:
        """ChaChaKey::eval (dpf.rs:103-142) at each index on the GPU: (n, 16) uint8."""
        idx = np.ascontiguousarray(np.atleast_1d(np.asarray(indices, dtype=np.uint64)))
        out = np.zeros((max(len(idx), 1), 16), dtype=np.uint8)
        kb = _blob(key)
        _check(lib.mg_key_eval(self._h, _u8(kb), len(key),
                               idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(idx),
                               _u8(out)))
        return out[:len(idx)]

    def eval_subtree(self, key: bytes, start: int, size: int) -> np.ndarray:
        """ChaChaKey::eval_subtree (dpf.rs:160-196) on the GPU: (size, 16) uint8."""
        out = np.zeros((max(size, 1), 16), dtype=np.uint8)
        kb = _blob(key)
        _check(lib.mg_key_eval_subtree(self._h, _u8(kb), len(key), start, size, _u8(out)))
        return out[:size]

    def full_eval(self, key: bytes) -> np.ndarray:
        """ChaChaKey::full_eval (dpf.rs:147-158) on the GPU: (2^d, 16) uint8."""
        d = parse_key(key).domain_bits
        out = np.zeros((1 << d, 16), dtype=np.uint8)
        kb = _blob(key)
        _check(lib.mg_key_full_eval(self._h, _u8(kb), len(key), _u8(out), out.nbytes))
        return out

    def scan_delta(self, delta: "DeltaBuffer", keys: Sequence[bytes]) -> np.ndarray:
        """scan_delta_for_gpu (scan.rs:434-485) for one leaf per row."""
        n = len(keys)
        out = np.zeros((max(n, 1), delta.row_bytes), dtype=np.uint8)
        blob, lens = _pack_keys(keys)
        _check(lib.mg_scan_delta(self._h, delta._h, _u8(blob), lens, n, _u8(out)))
        return out[:n]


class GpuPageMatrix:
    """GpuPageMatrix (storage.rs:18-145) with W-byte rows."""

    def __init__(self, scanner: GpuScanner, n_rows: int, row_bytes: int):
        h = ctypes.c_void_p()
        _check(lib.mg_matrix_create(scanner._h, n_rows, row_bytes, ctypes.byref(h)))
        self._h = h
        self.scanner = scanner  # the device must outlive this matrix
        self.n_rows = n_rows
        self.row_bytes = row_bytes

    @classmethod
    def alloc_empty(cls, scanner, n_rows, row_bytes):
        return cls(scanner, n_rows, row_bytes)

    @classmethod
    def from_file(cls, scanner, path: str, row_bytes: int, pad_tail: bool = True):
        """A raw row file streamed into a new device matrix (load_matrix_from_file +
        from_chunked_matrix, bench_page_gpu_batch.rs:326-379, storage.rs:46-86).
        pad_tail=False refuses a file that is not whole rows (snapshot_config.rs:152)."""
        h = ctypes.c_void_p()
        _check(lib.mg_matrix_load_file(scanner._h, os.fsencode(path), row_bytes,
                                       1 if pad_tail else 0, ctypes.byref(h)))
        m = cls.__new__(cls)
        m._h = h
        m.scanner = scanner
        m.row_bytes = row_bytes
        n = ctypes.c_uint64()
        w = ctypes.c_uint32()
        _check(lib.mg_matrix_info(h, ctypes.byref(n), ctypes.byref(w)))
        m.n_rows = n.value
        return m

    @classmethod
    def new(cls, scanner, rows: np.ndarray):
        rows = np.ascontiguousarray(rows, dtype=np.uint8)
        m = cls(scanner, rows.shape[0], rows.shape[1])
        m.update_rows(0, rows)
        return m

    def update_rows(self, start_row: int, rows: np.ndarray):
        rows = np.ascontiguousarray(rows, dtype=np.uint8)
        assert rows.ndim == 2 and rows.shape[1] == self.row_bytes
        _check(lib.mg_matrix_upload(self._h, start_row, rows.ctypes.data, rows.shape[0]))

    def download(self, start_row=0, n_rows=None) -> np.ndarray:
        n = self.n_rows - start_row if n_rows is None else n_rows
        out = np.zeros((n, self.row_bytes), dtype=np.uint8)
        _check(lib.mg_matrix_download(self._h, start_row, out.ctypes.data, n))
        return out

    def fill_seeded(self, seed: int):
        _check(lib.mg_matrix_fill_seeded(self._h, seed))

    def fill_with_pattern(self, seed: int):
        _check(lib.mg_matrix_fill_pattern(self._h, seed))

    def copy_from(self, other: "GpuPageMatrix"):
        _check(lib.mg_matrix_copy(self._h, other._h))

    def apply_delta(self, delta: "DeltaBuffer"):
        _check(lib.mg_matrix_apply_delta(self._h, delta._h))

    def num_pages(self):
        return self.n_rows

    def size_bytes(self):
        return self.n_rows * self.row_bytes

    def close(self):
        """Frees the rows; refused (MgError, MG_ERR_STATE) while a request loop serves them."""
        if self._h:
            _check(lib.mg_matrix_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeltaBuffer:
    """Device copy of a DeltaBuffer snapshot (delta.rs:55-307): XOR patches."""

    def __init__(self, scanner: GpuScanner, row_bytes: int, capacity: int):
        h = ctypes.c_void_p()
        _check(lib.mg_delta_create(scanner._h, row_bytes, capacity, ctypes.byref(h)))
        self._h = h
        self.scanner = scanner  # the device must outlive this buffer
        self.row_bytes = row_bytes
        self.capacity = capacity

    def set(self, row_idx: np.ndarray, diffs: np.ndarray, epoch: int = 0):
        idx = np.ascontiguousarray(row_idx, dtype=np.uint64)
        dif = np.ascontiguousarray(diffs, dtype=np.uint8).reshape(len(idx), self.row_bytes)
        _check(lib.mg_delta_set(self._h, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                _u8(dif), len(idx), epoch))

    def info(self):
        n, e = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib.mg_delta_info(self._h, ctypes.byref(n), ctypes.byref(e)))
        return n.value, e.value

    def close(self):
        if self._h:
            lib.mg_delta_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuServer:
    """Continuous-batching request loop over a device-resident matrix (C ABI
    mg_server_*). Requests of any key count are admitted as key slots free
    up and share every sweep step; answers equal a whole-matrix scan of the
    matrix state at admission (base rows XOR the updates applied before)."""

    def __init__(self, scanner: GpuScanner, matrix: GpuPageMatrix, segments=0, lane_depth=0,
                 max_active_keys=0, domain_bits=0, hbm_reserve_bytes=0, max_delta_entries=0,
                 double_buffer=False, deadline_ms=0, use_graph=False):
        o = ServerOpts(segments, lane_depth, max_active_keys, domain_bits, hbm_reserve_bytes,
                       max_delta_entries, 1 if double_buffer else 0, deadline_ms,
                       1 if use_graph else 0)
        h = ctypes.c_void_p()
        _check(lib.mg_server_create(scanner._h, matrix._h, ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self.scanner = scanner  # device and matrix must outlive the server
        self.matrix = matrix
        self.row_bytes = matrix.row_bytes

    def submit(self, keys: Sequence[bytes]) -> int:
        blob, lens = _pack_keys(keys)
        t = ctypes.c_uint64()
        _check(lib.mg_server_submit(self._h, _u8(blob), lens, len(keys), ctypes.byref(t)))
        return t.value

    def wait(self, ticket: int, n_keys: int, timeout_ms: int = 0):
        """(answers n_keys x W, info dict); MgError(MG_ERR_BUSY) on timeout."""
        out = np.zeros((max(n_keys, 1), self.row_bytes), dtype=np.uint8)
        info = RequestInfo()
        _check(lib.mg_server_wait(self._h, ticket, _u8(out), timeout_ms, ctypes.byref(info)))
        return out[:n_keys], info.as_dict()

    def abandon(self, ticket: int):
        """Gives up a ticket: a queued request never runs, one in flight leaves the sweep at
        the next step, a finished one is dropped."""
        _check(lib.mg_server_abandon(self._h, ticket))

    def update(self, row_idx: np.ndarray, diffs: np.ndarray, epoch: int):
        idx = np.ascontiguousarray(row_idx, dtype=np.uint64)
        dif = np.ascontiguousarray(diffs, dtype=np.uint8).reshape(len(idx), self.row_bytes)
        _check(lib.mg_server_update(self._h, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                    _u8(dif), len(idx), epoch))

    def inject_failure(self, code: int):
        """Makes the loop's next turn fail with `code` (tests of the failure path)."""
        _check(lib.mg_server_inject_failure(self._h, code))

    def stats(self):
        st = ServerStats()
        _check(lib.mg_server_stats_get(self._h, ctypes.byref(st)))
        return st.as_dict()

    def submit_snapshot(self, matrix: GpuPageMatrix, epoch: int):
        """EpochManager::submit_snapshot (epoch.rs:407-463): serve `matrix` for every
        request submitted from now on, at `epoch`; pending updates are dropped.
        Returns once the swap is done; the previous matrix is then free to close."""
        _check(lib.mg_server_snapshot(self._h, matrix._h, epoch))
        self.matrix = matrix

    def close(self):
        if self._h:
            lib.mg_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------- wire API ----
WIRE_MAX_BATCH = 32        # dto.rs:83 MAX_BATCH_SIZE
WIRE_MAX_BODY = 64 * 1024  # api/mod.rs:849 MAX_REQUEST_BODY_SIZE
WIRE_PATH = "/query/page/gpu"
WIRE_BATCH_PATH = "/query/page/gpu/batch"


def _bytes_arg(x):
    if x is None:
        return None
    return x.encode() if isinstance(x, str) else bytes(x)


def wire_parse(body, path=WIRE_PATH, method="POST", content_type="application/json",
               max_body_bytes=0, max_batch=0):
    """Routing, extractor and handler validation of one request
    (page_query_gpu(_batch)_handler, api/mod.rs:544-842) without a device.
    Returns (http_status, keys, n_queries): keys is the flat list of
    3 x n_queries wire keys when the status is 200, else empty."""
    body = _bytes_arg(body) or b""
    o = WireOpts(max_body_bytes, max_batch, 0)
    cap_keys = 3 * max(max_batch or WIRE_MAX_BATCH, 1)
    keys_cap = max(len(body) // 2, 1)
    buf = np.zeros(keys_cap, dtype=np.uint8)
    lens = (ctypes.c_size_t * cap_keys)()
    nq = ctypes.c_uint32()
    st = ctypes.c_int()
    rc = lib.mg_wire_parse(ctypes.byref(o), _bytes_arg(method), _bytes_arg(path),
                           _bytes_arg(content_type), body, len(body), _u8(buf), keys_cap, lens,
                           cap_keys, ctypes.byref(nq), ctypes.byref(st))
    if rc != MG_OK:
        return st.value, [], 0
    keys, off = [], 0
    for i in range(3 * nq.value):
        keys.append(buf[off:off + lens[i]].tobytes())
        off += lens[i]
    return st.value, keys, nq.value


def wire_format(epoch: int, pages: np.ndarray, batch: bool) -> bytes:
    """JSON body of a response: pages is (n_queries, 3, W) or (3, W)."""
    p = np.ascontiguousarray(pages, dtype=np.uint8)
    if p.ndim == 2:
        p = p[None]
    nq, _, w = p.shape
    cap = lib.mg_wire_response_capacity(nq, w)
    out = ctypes.create_string_buffer(cap)
    n = ctypes.c_size_t()
    _check(lib.mg_wire_format(epoch, _u8(p), nq, w, 1 if batch else 0, out, cap, ctypes.byref(n)))
    return out.raw[:n.value]


class WireApi:
    """The GPU page-query endpoints (POST /query/page/gpu and
    /query/page/gpu/batch) over a GpuServer request loop; `handle` is the
    transport-free handler an HTTP layer calls (morphogenesis_amd.http_api)."""

    def __init__(self, server: Optional[GpuServer], max_body_bytes=0, max_batch=0,
                 timeout_ms=0):
        self.server = server
        self.opts = WireOpts(max_body_bytes, max_batch, timeout_ms)
        w = server.row_bytes if server is not None else 16
        self.cap = lib.mg_wire_response_capacity(max_batch or WIRE_MAX_BATCH, w)

    def handle(self, method: str, path: str, content_type: Optional[str], body: bytes):
        """(http_status, response body bytes, error message)."""
        body = bytes(body or b"")
        out = ctypes.create_string_buffer(self.cap)
        n = ctypes.c_size_t()
        st = ctypes.c_int()
        h = self.server._h if self.server is not None else None
        _check(lib.mg_wire_handle(h, ctypes.byref(self.opts), _bytes_arg(method),
                                  _bytes_arg(path), _bytes_arg(content_type), body, len(body),
                                  out, self.cap, ctypes.byref(n), ctypes.byref(st)))
        err = "" if st.value == 200 else (lib.mg_last_error() or b"").decode(errors="replace")
        return st.value, out.raw[:n.value], err


# ----------------------------------------------------------- client side ---
CUCKOO_SEEDS = (0x12345678, 0x9ABCDEF0, 0xFEDCBA98)  # CuckooAddresser::new (cuckoo.rs:23-25)
ROWS_PER_PAGE = 16      # crates/morphogen-dpf/src/page.rs:23
ROW_SIZE_BYTES = 256    # page.rs:24
PAGE_SIZE_BYTES = 4096  # page.rs:25


def _u64_arr(vals):
    a = np.ascontiguousarray(np.array(vals, dtype=np.uint64))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def _blob(data: bytes):
    return np.frombuffer(bytes(data) or b"\0", dtype=np.uint8).copy()


def cuckoo_hash_indices(key: bytes, num_rows: int, seeds=CUCKOO_SEEDS) -> List[int]:
    """CuckooAddresser::with_seeds(num_rows, seeds).hash_indices(key) (cuckoo.rs:25-40)."""
    buf = _blob(key)
    sd, sdp = _u64_arr(seeds)
    out, outp = _u64_arr([0, 0, 0])
    _check(lib.mg_cuckoo_hash_indices(_u8(buf), len(key), num_rows, sdp, outp))
    return [int(x) for x in out]


@dataclass(frozen=True)
class PageAddress:
    """PageAddress (crates/morphogen-dpf/src/page.rs:39-50)."""
    page_index: int
    row_offset: int

    @staticmethod
    def from_row_index(row_index: int) -> "PageAddress":
        return PageAddress(row_index // ROWS_PER_PAGE, row_index % ROWS_PER_PAGE)


@dataclass
class GpuPageQueryKeys:
    """GpuPageQueryKeys (crates/morphogen-client/src/gpu.rs:11-15): wire keys."""
    page_addresses: List[PageAddress]
    keys_a: List[bytes]
    keys_b: List[bytes]


def generate_gpu_page_query(account_key: bytes, num_pages: int, seeds=CUCKOO_SEEDS,
                            domain_bits: int = 0, rng_seed: Optional[int] = None
                            ) -> GpuPageQueryKeys:
    """generate_gpu_page_query (gpu.rs:17-49). domain_bits 0 = ceil(log2 num_pages);
    rng_seed None draws the root seeds' seed from the OS RNG."""
    if not domain_bits:
        domain_bits = max(1, (max(num_pages, 1) - 1).bit_length())
    if rng_seed is None:
        rng_seed = int.from_bytes(os.urandom(8), "little")
    buf = _blob(account_key)
    sd, sdp = _u64_arr(seeds)
    klen = 19 + 18 * domain_bits + 16 if 0 < domain_bits <= 32 else 1
    ka = np.zeros(3 * klen, dtype=np.uint8)
    kb = np.zeros(3 * klen, dtype=np.uint8)
    pidx, pidxp = _u64_arr([0, 0, 0])
    offs = (ctypes.c_uint32 * 3)()
    rc = lib.mg_page_query(_u8(buf), len(account_key), num_pages, sdp, domain_bits,
                           rng_seed & (2**64 - 1), _u8(ka), _u8(kb), pidxp, offs)
    if rc != MG_OK:
        raise _ERRORS.get(rc, MgError)(rc, f"page query over {num_pages} pages, d={domain_bits}")
    return GpuPageQueryKeys(
        [PageAddress(int(pidx[i]), int(offs[i])) for i in range(3)],
        [ka[i * klen:(i + 1) * klen].tobytes() for i in range(3)],
        [kb[i * klen:(i + 1) * klen].tobytes() for i in range(3)])


class PageAggregationError(MgError):
    """EpochMismatch / PageLengthMismatch (crates/morphogen-client/src/page.rs:54-65)."""


class ExtractError(MgError):
    """InvalidRowOffset (page.rs:163-166)."""


def aggregate_gpu_page_responses(epoch_a: int, pages_a: Sequence[bytes], epoch_b: int,
                                 pages_b: Sequence[bytes]):
    """aggregate_page_responses (page.rs:67-103): (epoch, [a_i XOR b_i])."""
    if len(pages_a) != len(pages_b):
        raise PageAggregationError(MG_ERR_INVALID_ARG, "page count mismatch")
    out = []
    for i, (a, b) in enumerate(zip(pages_a, pages_b)):
        if len(a) != len(b):
            raise PageAggregationError(MG_ERR_INVALID_ARG,
                                       f"page {i}: lengths {len(a)} and {len(b)}")
        ba, bb = _blob(a), _blob(b)
        o = np.zeros(max(len(a), 1), dtype=np.uint8)
        rc = lib.mg_page_aggregate(epoch_a, _u8(ba), epoch_b, _u8(bb), len(a), 1, _u8(o))
        if rc == MG_ERR_STATE:
            raise PageAggregationError(rc, f"epoch mismatch: {epoch_a} vs {epoch_b}")
        _check(rc)
        out.append(o[:len(a)].tobytes())
    return epoch_a, out


def extract_row_from_page(page: bytes, row_offset: int) -> bytes:
    """extract_row_from_page (crates/morphogen-dpf/src/page.rs:698-708)."""
    buf = _blob(page)
    out = np.zeros(ROW_SIZE_BYTES, dtype=np.uint8)
    rc = lib.mg_page_extract_row(_u8(buf), len(page), row_offset, _u8(out))
    if rc == MG_ERR_ROWS:
        raise ExtractError(rc, f"row offset {row_offset}")
    _check(rc)
    return out.tobytes()


def extract_rows_from_gpu_pages(pages: Sequence[bytes], addresses: Sequence[PageAddress]):
    """extract_rows_from_pages (page.rs:140-162)."""
    return [extract_row_from_page(p, a.row_offset) for p, a in zip(pages, addresses)]


def page_query_body(keys: Sequence[bytes]) -> bytes:
    """Request body of POST /query/page/gpu (dto.rs:126-130)."""
    return ('{"keys":[' + ",".join('"0x' + bytes(k).hex() + '"' for k in keys) + "]}").encode()


def parse_page_response(body: bytes):
    """(epoch_id, pages) of a /query/page/gpu response (dto.rs PageQueryResponse)."""
    import json
    doc = json.loads(body)
    return doc["epoch_id"], [bytes.fromhex(p[2:] if p.startswith("0x") else p)
                             for p in doc["pages"]]

