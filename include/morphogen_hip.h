/*
 * morphogen_hip.h - C ABI of the MI355X (gfx950) fused ChaCha8-DPF
 * masked-XOR PIR scan.
 *
 * Plain C: opaque handles, pointers and sizes, int status codes, caller-owned
 * host buffers. No call aborts or throws across this boundary; on failure a
 * non-zero MG_* code is returned and mg_last_error() describes it (per
 * calling thread).
 *
 * Threading (the reference shares one GpuScanner across threads with a
 * ticket-based pool of workspace slots, kernel.rs:50-72, 733-739): every
 * entry point is safe to call from many threads. Scans and key evaluations
 * take one of the device's scan workspaces (a HIP stream with its own
 * buffers; mg_device_set_workspaces) and run concurrently with each other on
 * distinct streams. Writes to a matrix or delta (upload, fill, copy, apply,
 * set, destroy) wait for the scans in flight and hold new ones back while
 * they run, so a scan never sees a matrix half written.
 *
 * Each entry point names the reference interface it replaces
 * (igor53627/morphogenesis, paths relative to the repository root). The
 * reference's page leaves (4096 B) and 25-bit domain cap are generalised to
 * one DPF leaf per W-byte row (W a multiple of 16) and d <= 32.
 *
 * Answer semantics for one key (kernel.rs:1243-1256, 1374-1383):
 *   ans[j] = XOR over rows i < n_rows of rows[i][j] & mask(i)[j mod 16]
 * where mask(i) is the 16-byte DPF leaf of key at index i (dpf.rs:103-142).
 */
#ifndef MORPHOGEN_HIP_H
#define MORPHOGEN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MG_ABI_VERSION 1

/* Status codes. */
enum {
  MG_OK = 0,
  MG_ERR_INVALID_ARG = 1,   /* null handle/pointer, bad size or option      */
  MG_ERR_KEY_FORMAT = 2,    /* GpuDpfError::InvalidKeyFormat (dpf.rs:277-290) */
  MG_ERR_DOMAIN = 3,        /* d > 32, or keys of one call disagree on d     */
  MG_ERR_ROWS = 4,          /* n_rows > 2^d (kernel.rs:1283-1285), bad W,
                               row range or delta row index out of bounds    */
  MG_ERR_OUT_OF_MEMORY = 5, /* device allocation failed                     */
  MG_ERR_DEVICE = 6,        /* HIP runtime error                            */
  MG_ERR_NO_DEVICE = 7,     /* no such GPU ordinal                          */
  MG_ERR_BUSY = 8,          /* server queue full / admission refused        */
  MG_ERR_STATE = 9          /* call not valid in the object's current state */
};

typedef struct mg_device mg_device;   /* GpuScanner (kernel.rs:158-171)      */
typedef struct mg_matrix mg_matrix;   /* GpuPageMatrix (storage.rs:18-22)    */
typedef struct mg_delta mg_delta;     /* device copy of a DeltaBuffer
                                         snapshot (delta.rs:55-307)          */
typedef struct mg_server mg_server;   /* continuous-batching request loop    */

/* Parsed key header (ChaChaKey fields, dpf.rs:91-99). */
typedef struct {
  uint32_t domain_bits;
  uint32_t party;
  uint32_t root_t;
  uint32_t wire_len;
} mg_key_info;

/* Per-call timing (KernelTiming, kernel.rs:1213-1228), milliseconds, whole
 * call (the reference averages per query). kernel_ms is measured with HIP
 * events on the stream the scan kernels run on. */
typedef struct {
  double h2d_ms;
  double kernel_ms;
  double d2h_ms;
  double total_ms;
  double scan_kernel_ms; /* the fused scan kernel(s) alone                 */
  uint32_t kernel;       /* MG_KERNEL_* actually used                      */
  uint32_t lane_depth;   /* kl actually used                               */
  uint32_t keys_per_wave;
  uint32_t waves;        /* waves launched for the fused scan              */
  uint32_t stack_global; /* lane-DFS stack levels kept in global memory    */
  uint32_t dfs_waves;    /* split kernel: DFS waves per pair (else 0)     */
} mg_timing;

enum {
  MG_KERNEL_AUTO = 0,     /* lane-DFS up to 128 B rows; the split kernel
                             for 528-1024 B rows when a pass has >= 3 keys
                             or the rows are 1024 B (and d >= kl + 6); the
                             stream kernel for other rows up to 4096 B    */
  MG_KERNEL_LANE_DFS = 1, /* per-lane subtree DFS, rows of <= 128 B       */
  MG_KERNEL_STREAM = 2,   /* LDS-staged masks, coalesced wide rows (rows of
                             <= 512 B packed several to a load)           */
  MG_KERNEL_POINT = 3,    /* one full path walk per row                    */
  MG_KERNEL_SUFFIX = 4,   /* comparison baseline shaped like the reference
                             kernel (fused_kernel.cu:309-360): 2048-row
                             tiles, one prefix walk per tile, an 11-level
                             walk per row; rows of <= 128 B               */
  MG_KERNEL_STREAM_SPLIT = 5 /* the stream kernel with each block split
                             into a DFS wave and a row-streaming wave;
                             rows of >= 16 B and <= 4096 B, d >= kl + 6   */
};

/* Scan options; zero-initialise for defaults. The tuning fields replace the
 * reference's kernel flavour / tile flags of scan_batch_opts
 * (kernel.rs:711-720). */
typedef struct {
  uint32_t kernel;          /* MG_KERNEL_*                                 */
  uint32_t lane_depth;      /* kl: leaves per lane = 2^kl; 0 = tuned       */
  uint32_t keys_per_wave;   /* stream and split kernels, 1-4, clamped to   */
                            /* the width's cap (stream: 3 at 2 KiB rows,   */
                            /* 2 at 4 KiB; split: 3 at 4 KiB, 4 in column  */
                            /* halves); 0 = tuned                          */
  uint32_t waves_per_key;   /* 0 = tuned                                   */
  uint32_t waves_per_block; /* lane-DFS, stream 1-4; split 2, 4, 8;
                               0 = tuned                                   */
  uint32_t use_graph;       /* replay the launch sequence as a HIP graph   */
  uint64_t row_begin;       /* scan rows [row_begin, row_end) only; the    */
  uint64_t row_end;         /* range must be tile-aligned unless it is the */
                            /* whole matrix; row_end 0 = n_rows            */
  const mg_delta *delta;    /* if set: answers ^= delta contribution
                               (scan_delta_for_gpu, scan.rs:434-485)       */
  uint32_t stack_global;    /* lane-DFS: 0 = tuned, else 1 + the number of
                               top stack levels kept in global memory
                               instead of LDS (0-4; 1 = none)              */
  uint32_t row_loads;       /* lane-DFS: 0 = tuned, 1 = a leaf pair's rows
                               are loaded before its expansions, 2 = after */
  uint32_t row_halves;      /* split kernel, rows over 1 KiB: 0 = tuned
                               (halves over 2 KiB from 4 keys), 1 = one
                               pair per row, 2 = two pairs, one per column
                               half of 64 (to 2 KiB) or 128 chunks          */
  uint32_t mask_depth;      /* split kernel: leaf-pair steps of masks a DFS
                               wave may run ahead of its streaming wave,
                               2-8; 0 = tuned (as many as LDS holds, <= 4)  */
} mg_scan_opts;

int mg_abi_version(void);
const char *mg_last_error(void);

/* GpuScanner::new(device_ord) (kernel.rs:258). */
int mg_device_open(int ordinal, mg_device **out);
/* MG_ERR_STATE (nothing freed) while matrices, deltas or request loops
 * opened on the device are still open. */
int mg_device_close(mg_device *dev);
/* Scan workspaces the device's calls spread over, 1-8 (default 4: the
 * reference pool's default and cap, kernel.rs:50-65). A call takes a free
 * one, or waits for one when all are busy; mg_scan_per_query deals its
 * queries over the first n. Workspaces beyond n keep their buffers. */
#define MG_MAX_WORKSPACES 8u
int mg_device_set_workspaces(mg_device *dev, uint32_t n);
/* Free/total HBM and compute-unit count of the device. */
int mg_device_info(mg_device *dev, uint64_t *free_bytes, uint64_t *total_bytes,
                   uint32_t *compute_units, char *name, size_t name_len);

/* GpuPageMatrix::alloc_empty (storage.rs:89-101): zeroed rows. */
int mg_matrix_create(mg_device *dev, uint64_t n_rows, uint32_t row_bytes,
                     mg_matrix **out);
/* GpuPageMatrix::update_pages (storage.rs:104-124) / new (storage.rs:27-43):
 * copy n_rows host rows to rows [first_row, first_row + n_rows). */
int mg_matrix_upload(mg_matrix *m, uint64_t first_row, const void *rows,
                     uint64_t n_rows);
int mg_matrix_download(mg_matrix *m, uint64_t first_row, void *rows,
                       uint64_t n_rows);
/* Seeded synthetic rows generated on the device: 64-bit word j of the
 * row-major matrix = splitmix64 output j for `seed` (little-endian). */
int mg_matrix_fill_seeded(mg_matrix *m, uint64_t seed);
/* ChunkedMatrix::fill_with_pattern (chunked_matrix.rs:83-93), generated on
 * the device with LCG jump-ahead. */
int mg_matrix_fill_pattern(mg_matrix *m, uint64_t seed);
int mg_matrix_info(const mg_matrix *m, uint64_t *n_rows, uint32_t *row_bytes);
/* Device-to-device copy of all rows between two matrices of one shape (a
 * snapshot for rotation, epoch.rs:407-463 submit_snapshot). */
int mg_matrix_copy(mg_matrix *dst, const mg_matrix *src);
/* MG_ERR_STATE (nothing freed) while a request loop serves the matrix or
 * holds it as a queued snapshot. */
int mg_matrix_destroy(mg_matrix *m);

/* Raw row file -> new device matrix (load_matrix_from_file,
 * crates/morphogen-server/src/bin/bench_page_gpu_batch.rs:326-379, then
 * GpuPageMatrix::from_chunked_matrix, crates/morphogen-gpu-dpf/src/
 * storage.rs:46-86). The file is rows back to back (e.g. the 32-byte Compact
 * schema, docs/DATA_STRUCTURES.md:17-26). With MG_LOAD_PAD_TAIL a ragged last
 * row is zero-padded, as the reference pads to a whole page; without it the
 * file must hold whole rows (decode_snapshot_matrix, network/
 * snapshot_config.rs:144-159). An empty file is MG_ERR_ROWS. Reads overlap
 * the copies (two pinned 64 MiB staging buffers on a stream of its own), so
 * a loop can keep serving while the next snapshot loads. */
#define MG_LOAD_PAD_TAIL 1u
int mg_matrix_load_file(mg_device *dev, const char *path, uint32_t row_bytes, uint32_t flags,
                        mg_matrix **out);

/* ChaChaKey::from_bytes (dpf.rs:277-316): exact length 19 + 18d + 16, plus
 * the evaluator's limit d <= 32 (ChaChaParams::new, dpf.rs:65-78). */
int mg_key_parse(const uint8_t *bytes, size_t len, mg_key_info *out);

/* ChaChaKey::eval (dpf.rs:103-142) on the GPU: out[16 i .. 16 i + 16) = the
 * key's 16-byte leaf (Seed128, four little-endian words) at indices[i], for
 * i < n. Only the low d bits of an index are read, as in the reference. */
int mg_key_eval(mg_device *dev, const uint8_t *key, size_t key_len,
                const uint64_t *indices, uint64_t n, uint8_t *out);
/* ChaChaKey::eval_subtree (dpf.rs:160-196) on the GPU: the leaves of
 * [start, start + size) in order, 16 bytes each. size must be a power of two
 * and start a multiple of it (MG_ERR_KEY_FORMAT, the reference's
 * InvalidKeyFormat); a subtree larger than the 2^d-leaf domain is MG_ERR_ROWS
 * (the reference writes a partial pattern there). */
int mg_key_eval_subtree(mg_device *dev, const uint8_t *key, size_t key_len,
                        uint64_t start, uint64_t size, uint8_t *out);
/* ChaChaKey::full_eval (dpf.rs:147-158) on the GPU: all 2^d leaves; out_len
 * must be exactly 2^d x 16 bytes (MG_ERR_INVALID_ARG, InvalidKeyLength). */
int mg_key_full_eval(mg_device *dev, const uint8_t *key, size_t key_len, uint8_t *out,
                     uint64_t out_len);

/* generate_chacha_dpf_keys (dpf.rs:322-432), client side, host only: writes
 * the two parties' wire keys (19 + 18d + 16 bytes each) for a point function
 * at `target` over a d-bit domain. Root seeds are derived from `seed`
 * (splitmix64 words 0-1 and 2-3) instead of the OS RNG. */
int mg_keygen(uint32_t domain_bits, uint64_t target, uint64_t seed,
              uint8_t *key0, uint8_t *key1);

/* CuckooAddresser::with_seeds(num_rows, seeds).hash_indices(key)
 * (crates/morphogen-core/src/cuckoo.rs:25-40): out[i] = SipHash-1-3 with the
 * all-zero key (Rust's DefaultHasher) over LE64(seeds[i]) || key, modulo
 * num_rows. CuckooAddresser::new's seeds are 0x12345678, 0x9ABCDEF0,
 * 0xFEDCBA98. num_rows = 0: MG_ERR_ROWS. */
int mg_cuckoo_hash_indices(const uint8_t *key, size_t key_len, uint64_t num_rows,
                           const uint64_t seeds[3], uint64_t out[3]);

#define MG_ROWS_PER_PAGE 16    /* crates/morphogen-dpf/src/page.rs:23 */
#define MG_PAGE_ROW_BYTES 256  /* page.rs:24 */
#define MG_PAGE_BYTES 4096     /* page.rs:25 */

/* generate_gpu_page_query (crates/morphogen-client/src/gpu.rs:17-49): the
 * account's 3 cuckoo rows among num_pages x 16, their page addresses
 * (page = row / 16, offset = row % 16) and a key pair per page; keys_a and
 * keys_b receive 3 wire keys each, 19 + 18d + 16 bytes apart. Root seeds
 * derive from rng_seed (independent per pair) instead of the OS RNG.
 * MG_ERR_DOMAIN for d outside 1..32, MG_ERR_ROWS when a page index does not
 * fit the domain (PageIndexTooLarge). */
int mg_page_query(const uint8_t *account_key, size_t key_len, uint64_t num_pages,
                  const uint64_t seeds[3], uint32_t domain_bits, uint64_t rng_seed,
                  uint8_t *keys_a, uint8_t *keys_b, uint64_t page_index[3],
                  uint32_t row_offset[3]);
/* aggregate_page_responses (crates/morphogen-client/src/page.rs:67-103):
 * out = pages_a XOR pages_b (n_pages pages of page_bytes); MG_ERR_STATE when
 * the two servers answered for different epochs. */
int mg_page_aggregate(uint64_t epoch_a, const uint8_t *pages_a, uint64_t epoch_b,
                      const uint8_t *pages_b, size_t page_bytes, uint32_t n_pages,
                      uint8_t *out);
/* extract_row_from_page (crates/morphogen-dpf/src/page.rs:698-708): the
 * 256-byte row at row_offset; MG_ERR_ROWS if row_offset >= 16 or the row
 * runs past the page. */
int mg_page_extract_row(const uint8_t *page, size_t page_bytes, uint32_t row_offset,
                        uint8_t *out);

/* GpuScanner::scan / scan_batch / scan_batch_optimized(_tiled)(_with_graph)
 * (kernel.rs:685-1162), flattened: n_keys keys of the same domain in one
 * pass over the matrix. keys = concatenated wire keys, key_lens[i] their
 * lengths. out = n_keys x row_bytes, caller-owned. timing may be NULL. */
int mg_scan(mg_device *dev, mg_matrix *m, const uint8_t *keys,
            const size_t *key_lens, uint32_t n_keys, uint8_t *out,
            const mg_scan_opts *opts, mg_timing *timing);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mg_scan with the same keys, options and answers, but
 * every device buffer the scan needs beyond the matrix and the delta (keys,
 * per-wave partials, lane-DFS stack levels, answers) lives in one fixed
 * array of MG_NOALLOC_SCRATCH_BYTES that the library holds on each GPU until
 * the process ends; keys are staged in host memory reserved at
 * mg_device_open. The call allocates nothing. A scan that needs more than
 * the array, or more than MG_NOALLOC_MAX_KEYS keys, is refused with
 * MG_ERR_INVALID_ARG; use_graph is refused too (instantiating a graph
 * allocates). Calls through one mg_device serialise on its control stream's
 * lock (and with matrix and delta writes); calls through two handles opened
 * on the same GPU share the array and race. */
#define MG_NOALLOC_SCRATCH_BYTES (64u << 20)
#define MG_NOALLOC_MAX_KEYS 1024u
int mg_scan_noalloc(mg_device *dev, mg_matrix *m, const uint8_t *keys,
                    const size_t *key_lens, uint32_t n_keys, uint8_t *out,
                    const mg_scan_opts *opts, mg_timing *timing);

/* scan_batch_single_query_multistream_optimized(_with_graph)
 * (kernel.rs:955-1097): one pass per query of keys_per_query keys, the
 * reference's "plain" serving mode; same answers as mg_scan. The queries are
 * dealt round-robin over the device's workspaces, each on its own stream
 * (the reference's forked streams, kernel.rs:1024-1037), so their launches,
 * key uploads and answer downloads overlap. timing: kernel, copy and scan
 * times summed over the queries, total_ms the call's wall time. */
int mg_scan_per_query(mg_device *dev, mg_matrix *m, const uint8_t *keys,
                      const size_t *key_lens, uint32_t n_keys,
                      uint32_t keys_per_query, uint8_t *out,
                      const mg_scan_opts *opts, mg_timing *timing);

/* DeltaBuffer snapshot on the device (delta.rs:55-307): entries
 * (row_idx[i], diffs[i * row_bytes ...]) XOR-patch rows. */
int mg_delta_create(mg_device *dev, uint32_t row_bytes, uint64_t capacity,
                    mg_delta **out);
int mg_delta_set(mg_delta *dl, const uint64_t *row_idx, const uint8_t *diffs,
                 uint64_t n_entries, uint64_t epoch);
int mg_delta_info(const mg_delta *dl, uint64_t *n_entries, uint64_t *epoch);
int mg_delta_destroy(mg_delta *dl);
/* scan_delta_for_gpu (scan.rs:434-485), one leaf per row: out =
 * n_keys x row_bytes answers of the delta entries alone. */
int mg_scan_delta(mg_device *dev, const mg_delta *dl, const uint8_t *keys,
                  const size_t *key_lens, uint32_t n_keys, uint8_t *out);
/* Epoch merge (epoch.rs:118-181 XOR patch + update_pages): rows[idx] ^= diff
 * for every entry; repeated indices compose by XOR. */
int mg_matrix_apply_delta(mg_matrix *m, const mg_delta *dl);

/* ---------------------------------------------------------------------------
 * Continuous-batching request loop (no reference counterpart as one object:
 * it replaces the server's GPU batch policy and dispatch,
 * crates/morphogen-server/src/network/gpu_batch.rs:129-149 and gpu_scan.rs:36-76,
 * plus the epoch-consistent scan, scan.rs:260-299).
 *
 * A host thread sweeps the matrix segment by segment. Each sweep step runs
 * the fused scan over one segment for every active key at once. A request is
 * admitted as soon as enough key slots are free, starts at the next segment,
 * and completes when its keys have seen every segment; XOR is order-free, so
 * its answers equal a whole-matrix scan. Each key answers for the matrix
 * state (base rows XOR every update applied before its admission), so
 * updates never tear an in-flight answer.
 * ------------------------------------------------------------------------- */
typedef struct {
  uint32_t segments;          /* sweep steps per full pass; 0 = one per 2^25
                                 rows, between 1 and 64                     */
  uint32_t lane_depth;        /* kl of the fused kernel; 0 = 10 for rows of
                                 at most 128 B, else 7                      */
  uint32_t max_active_keys;   /* 0 = bounded by free HBM (at most 32768)    */
  uint32_t domain_bits;       /* key domain served; 0 = ceil(log2 n_rows)   */
  uint64_t hbm_reserve_bytes; /* HBM left free for other allocations       */
  uint64_t max_delta_entries; /* update rows kept before a rebase; 0 = 2^20 */
  uint32_t double_buffer;     /* keep a second matrix copy so a rebase never
                                 waits for in-flight keys                   */
  uint32_t deadline_ms;       /* admission deadline: mg_server_submit refuses
                                 (MG_ERR_BUSY) a request that, by the loop's
                                 step-time model (fixed + per key, fitted to
                                 the steps run) and the requests queued and in
                                 flight, would finish later than this after
                                 its submit, or would make an earlier request
                                 miss its own; 0 = none (the reference answers
                                 such requests with 503, network/api/mod.rs:
                                 638, 828)                                  */
  uint32_t use_graph;         /* replay each sweep step's kernels (scan and
                                 reductions into the accumulators) from a HIP
                                 graph cached per segment                   */
} mg_server_opts;

typedef struct {
  double queue_ms;    /* submit -> admitted                                 */
  double service_ms;  /* admitted -> answers on the host                    */
  double latency_ms;  /* submit -> answers on the host                      */
  uint64_t epoch;     /* update epoch the answers reflect                   */
  uint32_t n_keys;
} mg_request_info;

typedef struct {
  uint64_t submitted, completed, keys_completed, sweep_steps, epoch;
  uint64_t delta_entries;     /* updates not yet merged into a base copy   */
  uint64_t rebases;
  uint32_t active_keys, queued_requests, max_active_keys, segments;
  double busy_ms;             /* time spent in sweep steps                 */
  uint64_t snapshots;         /* matrices swapped in by mg_server_snapshot */
  uint64_t n_rows;            /* rows of the matrix served now             */
  uint64_t rejected;          /* submits refused by the admission deadline */
  uint64_t abandoned;         /* requests given up before completing: while
                                 queued (never run) or in flight (their keys
                                 leave the sweep at the next step)          */
  double key_step_ms;         /* fitted sweep-step time per key, and ...   */
  double step_fixed_ms;       /* ... per step: the admission model         */
  uint64_t graph_launches;    /* sweep-step launches replayed from a graph */
} mg_server_stats;

int mg_server_create(mg_device *dev, mg_matrix *m, const mg_server_opts *opts,
                     mg_server **out);
/* Queues one request of n_keys keys (typically 3 cuckoo keys x k lookups).
 * MG_ERR_BUSY if it could never fit the key slots. */
int mg_server_submit(mg_server *s, const uint8_t *keys, const size_t *key_lens,
                     uint32_t n_keys, uint64_t *ticket);
/* Waits up to timeout_ms (0 = forever) for a request; out = n_keys x W.
 * MG_ERR_BUSY on timeout (the request stays queued). */
int mg_server_wait(mg_server *s, uint64_t ticket, uint8_t *out,
                   uint32_t timeout_ms, mg_request_info *info);
/* Gives up a ticket nobody will wait for (e.g. after a timed-out wait). A
 * request still queued is removed and never runs; one in flight leaves the
 * sweep at the next step, freeing its key slots; a finished one is dropped. */
int mg_server_abandon(mg_server *s, uint64_t ticket);
/* Epoch update: rows[row_idx[i]] ^= diffs[i] for every request admitted
 * after this call (EpochManager::try_advance, epoch.rs:465-549). Updates are
 * applied at the next sweep step, ahead of requests still waiting for key
 * slots (as the reference's merge worker advances epochs whatever queries
 * wait, epoch.rs:465-549, 685); those then answer at the newer epoch, which
 * mg_request_info reports. Keys already admitted keep their epoch. */
int mg_server_update(mg_server *s, const uint64_t *row_idx,
                     const uint8_t *diffs, uint64_t n_entries, uint64_t epoch);
/* Snapshot hot-swap (EpochManager::submit_snapshot, crates/morphogen-server/
 * src/epoch.rs:407-463, behind admin_snapshot_handler, network/api/mod.rs:
 * 190-237): the loop serves `m` (same row width and device, at most 2^d rows)
 * for every request submitted after this call, at epoch `epoch`. Requests
 * admitted before it finish on the old rows; updates not yet merged are
 * dropped, as the reference drains its pending delta. Blocks until the swap
 * is done; the old matrix is then no longer read and may be destroyed. The
 * loop owns the matrix it serves: merges write into it. */
int mg_server_snapshot(mg_server *s, mg_matrix *m, uint64_t epoch);
int mg_server_stats_get(mg_server *s, mg_server_stats *out);
/* Fault injection for tests of the failure path (no reference counterpart):
 * the loop's next turn fails with `code` (an MG_ERR_* value) as a device
 * error would: waiting requests and queued snapshots get the error, later
 * calls get MG_ERR_* "server stopped", and mg_server_destroy still works. */
int mg_server_inject_failure(mg_server *s, int code);
int mg_server_destroy(mg_server *s);

/* ---------------------------------------------------------------------------
 * Wire API of the GPU page-query endpoints, transport-free: the caller's
 * HTTP layer passes method, path, Content-Type and body and sends back the
 * status and JSON body. Replaces page_query_gpu_handler and
 * page_query_gpu_batch_handler (crates/morphogen-server/src/network/api/
 * mod.rs:544-842), their routes and body limit (mod.rs:849, 886-896), the
 * DTOs (dto.rs:83, 126-156) and the hex codec (serde_hex.rs:58-85):
 *   POST /query/page/gpu        {"keys":["0x..","0x..","0x.."]}
 *        -> {"epoch_id":E,"pages":["0x..","0x..","0x.."]}
 *   POST /query/page/gpu/batch  {"queries":[{"keys":[...]},...]}   (1..32)
 *        -> {"epoch_id":E,"results":[{"pages":[...]},...]}
 * Pages are row_bytes long (the reference's 4096-byte pages are rows of a
 * 4096-byte matrix). Status: 404/405 routing, 415 not JSON, 413 body over the
 * limit, 400 malformed JSON, 422 wrong shape or bad hex, 400 batch size, key
 * count or a key that does not parse (keys up to d = 32; the reference's
 * kernel stops at 25, kernel.rs:45), 503 request larger than the loop's key
 * slots or wait timed out, 500 device error.
 * ------------------------------------------------------------------------- */
#define MG_WIRE_KEYS_PER_QUERY 3   /* kernel.rs KEYS_PER_QUERY              */
#define MG_WIRE_MAX_BATCH 32       /* dto.rs:83 MAX_BATCH_SIZE              */
#define MG_WIRE_MAX_BODY 65536     /* api/mod.rs:849 MAX_REQUEST_BODY_SIZE  */

typedef struct {
  uint64_t max_body_bytes;  /* 0 = MG_WIRE_MAX_BODY                         */
  uint32_t max_batch;       /* 0 = MG_WIRE_MAX_BATCH                        */
  uint32_t timeout_ms;      /* wait for the loop; 0 = forever. A request that
                               times out (503) still runs to completion; its
                               answers are dropped then (mg_server_abandon). */
} mg_wire_opts;

/* Output bytes a response of n_queries queries of row_bytes pages needs. */
size_t mg_wire_response_capacity(uint32_t n_queries, uint32_t row_bytes);
/* Routing, extractor and handler checks only (no device): on MG_OK the
 * request is valid (*http_status = 200) and keys holds its 3 x n_queries
 * wire keys back to back (key_lens[i] each). Otherwise an MG_ERR_* code with
 * *http_status set to the status the endpoint answers. */
int mg_wire_parse(const mg_wire_opts *opts, const char *method, const char *path,
                  const char *content_type, const char *body, size_t body_len,
                  uint8_t *keys, size_t keys_cap, size_t *key_lens,
                  uint32_t key_lens_cap, uint32_t *n_queries, int *http_status);
/* JSON response body for answers pages[n_queries][3][row_bytes]; batch = 0
 * for the single-query shape. NUL-terminated; *out_len excludes the NUL. */
int mg_wire_format(uint64_t epoch, const uint8_t *pages, uint32_t n_queries,
                   uint32_t row_bytes, int batch, char *out, size_t out_cap,
                   size_t *out_len);
/* One request end to end through the request loop `s` (all keys of a batch
 * as one loop request, so they share one admission epoch). Returns MG_OK
 * whenever a response was produced, whatever its status; the body is empty
 * for errors (mg_last_error() says why). s = NULL answers valid requests
 * with 503. Safe to call from many threads on one loop. */
int mg_wire_handle(mg_server *s, const mg_wire_opts *opts, const char *method,
                   const char *path, const char *content_type, const char *body,
                   size_t body_len, char *out, size_t out_cap, size_t *out_len,
                   int *http_status);

#ifdef __cplusplus
}
#endif

#endif
