/*
 * gmpt_amd.h -- C-ABI of the MI355X GPU Merkle Patricia Trie library
 * (libgmpt_amd.so).  Plain pointers and sizes only.
 *
 * The reference exposes its engine as the C++ class GpuMPT::Compress::MPT
 * (mpt-with-compress/include/mpt/gpu_mpt.cuh:18-192) and a thin C-ABI for
 * go-ethereum (mpt-with-compress/lib/libgmpt.h:1-55, declared here in
 * libgmpt.h).  This header is the C-ABI that both sit on: the C++ class in
 * gpu_mpt.hpp calls only these functions.  Each entry point names the
 * reference method it replaces.
 *
 * Conventions
 *  - Return value: 0 on success, negative on error; gmpt_last_error() then
 *    describes the error (per calling thread).  Nothing is printed.
 *  - A forest holds num_tries independent tries on one device (a single trie
 *    of the reference is a forest of one).  Calls on one forest are not
 *    thread-safe; different forests may be used from different threads.
 *  - Batches use the reference's packed layout (utils.cuh:254-309): hex keys
 *    (one nibble per byte, terminator 16 last) with inclusive int32
 *    [start, end] pairs; values with inclusive int64 pairs.
 *  - Duplicate keys in one batch: the last occurrence wins (the CPU MPT's
 *    sequential semantics, cpu_mpt.cuh:359-378); an existing key is
 *    overwritten by both PhaseNU and LockNU.
 */
#ifndef GMPT_AMD_H
#define GMPT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gmpt_forest gmpt_forest;

typedef struct gmpt_batch {
  const uint8_t *keys_hexs;       /* hex nibbles of all keys */
  const int *keys_indexs;         /* 2 per key, inclusive */
  const uint8_t *values_bytes;    /* all values */
  const int64_t *values_indexs;   /* 2 per value, inclusive */
  const uint8_t *const *values_hps; /* optional caller host pointer per value, returned by gets */
  const int32_t *trie_ids;        /* optional trie per key (default: trie 0) */
  int n;                          /* number of keys */
  int64_t keys_bytes;             /* device batches: bytes in keys_hexs */
  int64_t values_bytes_total;     /* device batches: bytes in values_bytes */
  int on_device;                  /* 1: every pointer above is a device pointer */
  int adopt_values;               /* device batches: trie keeps pointers into values_bytes (no copy) */
} gmpt_batch;

typedef struct gmpt_stats {
  double last_hash_ms;            /* device time of the last hash pass, from its bins to its counters read */
  uint32_t hash_launches;         /* hash kernel launches (levels) of the last hash */
  uint32_t hashed_nodes;          /* nodes re-hashed by the last hash */
  uint32_t restarts;              /* LockNU optimistic restarts in the last insert */
  uint32_t lock_fails;            /* LockNU keys that restarted at least once */
  uint32_t fallback_sorts;        /* PhaseNU full-key re-sorts in the last insert */
  uint32_t regroups;              /* PhaseNU regroupings of split groups in the last insert */
  uint64_t perms;                 /* Keccak-f permutations run by the last hash */
  uint64_t enc_bytes;             /* node encoding bytes produced by the last hash */
  uint64_t pool_bytes_used;       /* HBM held by nodes, keys and values */
  uint64_t pool_bytes_reserved;   /* HBM reserved by the pools */
  double leaf_perm_ms;            /* device time of the leaf level's permutation kernel (last hash) */
  uint64_t leaf_perms;            /* its permutations */
  uint64_t leaf_bytes;            /* encoding bytes of the leaf level */
  uint32_t leaf_nodes;            /* leaf nodes hashed */
  uint32_t ext_direct_nodes;      /* extension nodes hashed by the leaf word builder (last hash) */
} gmpt_stats;

/* Creates a forest of num_tries empty tries on device `device`.
 * Replaces: GpuMPT::Compress::MPT::MPT() (gpu_mpt.cuh:162-173). */
int gmpt_forest_create(int device, uint32_t num_tries, gmpt_forest **out);
/* Replaces: MPT::~MPT() (gpu_mpt.cuh:186-191). */
void gmpt_forest_destroy(gmpt_forest *f);
const char *gmpt_last_error(void);

/* PhaseNU batch insert.  *targets / *n_targets receive the dirty-node list to
 * pass to gmpt_hash_onepass.
 * Replaces: MPT::puts_2phase_with_valuehp (gpu_mpt.cuh:95-97, 1262-1399). */
int gmpt_puts_2phase(gmpt_forest *f, const gmpt_batch *b, void **targets, int *n_targets);
/* LockNU batch insert.
 * Replaces: MPT::puts_latching_with_valuehp_v2 (gpu_mpt.cuh:69-71, 554-615). */
int gmpt_puts_latching(gmpt_forest *f, const gmpt_batch *b, void **targets, int *n_targets);
/* Mixed read/write batch: rw_flags[i] = 0 reads key i, 1 writes it (a host
 * array for host batches, a device array for device batches).  Reads see the
 * tries as they were before the batch (reads are ordered before the batch's
 * writes); *read_num results come back in batch order: the host pointer given
 * at insert and the value size (NULL and 0 for an absent key), into arrays of
 * capacity n.  The writes then run as one PhaseNU (mode 0) or LockNU (mode 1)
 * insert; *targets / *n_targets as for the plain inserts.
 * Replaces: MPT::puts_2phase_with_valuehp_with_read (gpu_mpt.cuh:105-110,
 * 1609-1713) and MPT::puts_latching_with_valuehp_v2_with_read
 * (gpu_mpt.cuh:74-78, 684-765); their experiment-recorder arguments have no
 * counterpart.  Read results there come back in arbitrary order. */
int gmpt_puts_with_read(gmpt_forest *f, const gmpt_batch *b, const uint8_t *rw_flags, int mode, int *read_num,
                        const uint8_t **read_values_hps, int *read_values_sizes, void **targets, int *n_targets);
/* PhaseHC: re-hash every dirty node bottom-up and refresh each trie's root.
 * Replaces: MPT::hash_onepass_v2 (gpu_mpt.cuh:131, 1162-1191). */
int gmpt_hash_onepass(gmpt_forest *f, void *targets, int n_targets);
/* Root hash of one trie: *size = 32, or the raw root encoding length when it
 * is shorter than 32 bytes, or 0 for an empty trie.
 * Replaces: MPT::get_root_hash (gpu_mpt.cuh:151-152, 1217-1247). */
int gmpt_get_root_hash(gmpt_forest *f, uint32_t trie, uint8_t out[32], int *size);
/* Root hashes of tries [first, first + count) with one device copy: out holds
 * 32 bytes per trie (zero padded), sizes[i] as for gmpt_get_root_hash.
 * Replaces: MPT::get_root_hash called once per trie (gpu_mpt.cuh:1217-1247). */
int gmpt_get_root_hashes(gmpt_forest *f, uint32_t first, uint32_t count, uint8_t *out, int *sizes);
/* For each key: the host pointer given at insert (values_hps) and the value
 * size; absent keys give NULL and size 0.
 * Replaces: MPT::gets_parallel (gpu_mpt.cuh:139-140, 984-1036). */
int gmpt_gets(gmpt_forest *f, const gmpt_batch *keys, const uint8_t **values_hps, int *values_sizes);
/* Like gmpt_gets but copies the stored value bytes to out (offsets[i] per key,
 * sizes[i] = -1 for absent keys).  No reference counterpart: the reference's
 * C-ABI inserts without host pointers and so cannot return values. */
int gmpt_get_values(gmpt_forest *f, const gmpt_batch *keys, uint8_t *out, int64_t out_cap, int64_t *offsets,
                    int *sizes);
/* Every node on the keys' paths once: 32-byte hash slots, RLP encodings with
 * inclusive u64 [start, end] pairs.  Arrays are released with gmpt_free_host.
 * Replaces: MPT::flush_dirty_nodes (gpu_mpt.cuh:146-149, 1039-1122). */
int gmpt_export_nodes(gmpt_forest *f, const gmpt_batch *keys, const uint8_t **hashes, const uint8_t **encs,
                      const uint64_t **enc_indexs, uint64_t *num);
void gmpt_free_host(const void *p);
int gmpt_reset_trie(gmpt_forest *f, uint32_t trie);
/* Empties every trie of the forest and recycles its pools without freeing
 * them.  Replaces constructing a fresh MPT per run (gpu_mpt.cuh:162-173),
 * which in the reference allocates a new 16 GB pool each time. */
int gmpt_forest_clear(gmpt_forest *f);
int gmpt_synchronize(gmpt_forest *f);
/* Tunables by name.  "hc_spread_max": PhaseHC levels with at most this many
 * dirty nodes run Keccak with one state spread over 25 lanes (latency), larger
 * levels one state per lane (throughput).  Default 1024, or the
 * GMPT_HC_SPREAD_MAX environment variable.  "hc_fuse_leaves" (0 or 1,
 * default 1): leaves whose key remainder has at most 64 nibbles and whose
 * value is under 16,383 bytes are encoded and hashed by one kernel, the
 * message built in registers, instead of an encode kernel, an HBM scratch
 * buffer and a permutation kernel.  "hc_fuse_inner" (0 or 1, default 1):
 * inner levels of mostly 1-block branches and extensions are encoded and
 * hashed in one kernel with the messages in LDS.  No reference counterpart.
 * "pool_memory" (0 device, 1 host): 1 makes the trie state (node pools, key
 * and value arenas, dirty list) grow in pinned host memory the kernels reach
 * over the host link, so a forest can hold tries larger than the GPU's
 * memory; replaces the unified-memory UMMPT (um_mpt.cuh:13-237). */
int gmpt_set_option(gmpt_forest *f, const char *name, int64_t value);
int gmpt_get_option(gmpt_forest *f, const char *name, int64_t *value);
/* Autotuner: inserts `sample` (mode 0 PhaseNU, 1 LockNU) into a scratch forest
 * on the same device, times PhaseHC for each candidate hc_spread_max, and sets
 * the fastest on f.  No reference counterpart. */
int gmpt_autotune(gmpt_forest *f, const gmpt_batch *sample, int mode, int64_t *chosen, double *best_ms);
int gmpt_get_stats(gmpt_forest *f, gmpt_stats *out);
/* Bulk rebuild of device batch b (on_device and adopt_values set; trie_ids
 * for a multi-trie forest) into the forest after emptying it: clear, PhaseNU
 * insert, PhaseHC hash, as one replay of a captured HIP graph.  The first call
 * for a batch -- its device buffers and sizes -- runs the pass eagerly,
 * captures it and replays it once; later calls with the same buffers (new
 * contents of the same shape) replay it with one host synchronisation
 * instead of four.  The device checks each replay against the captured pass
 * (no fallback sort, the same dirty-node count and hash levels); a batch that
 * differs is redone eagerly and captured again, so results are always those
 * of the plain calls.  *replayed = 1 when the replay stood.  Roots are then
 * read with gmpt_get_root_hash(es).  Replaces the batched form's sequence of
 * MPT calls per trie (gpu_mpt.cuh:1262-1399, 1162-1191); no reference
 * counterpart for the graph itself. */
int gmpt_rebuild_graph(gmpt_forest *f, const gmpt_batch *b, int *replayed);

/* Device helpers so callers can keep batches resident in HBM. */
int gmpt_device_count(int *count);
int gmpt_device_malloc(void **p, size_t bytes);
int gmpt_device_free(void *p);
int gmpt_memcpy_htod(void *dst, const void *src, size_t bytes);
int gmpt_memcpy_dtoh(void *dst, const void *src, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (16-byte aligned sizes);
 * also the known-traffic kernel that calibrates memory counters. */
int gmpt_device_copy(void *dst, const void *src, size_t bytes);
/* Device-time stopwatch on the forest's stream. */
int gmpt_timer_start(gmpt_forest *f);
int gmpt_timer_stop(gmpt_forest *f, float *ms);
/* Keccak-256 of n host messages on the device (test hook for the sponge).
 * Replaces: GPUHashMultiThread batch hashing (batch_mode_hash.cuh:89-115). */
int gmpt_keccak256_batch(const uint8_t *data, const uint64_t *offsets, const uint32_t *lengths, uint32_t n,
                         uint8_t *out);
/* As gmpt_keccak256_batch with the permutation layout chosen: mode 0 keeps one
 * state per GPU lane (PhaseHC's large levels), mode 1 spreads one state over 25
 * lanes (PhaseHC's small levels).  No reference counterpart (test hook). */
int gmpt_keccak256_batch_mode(const uint8_t *data, const uint64_t *offsets, const uint32_t *lengths, uint32_t n,
                              uint8_t *out, int mode);

/* Root hash of the trie that batch b builds from empty (mode 0: PhaseNU,
 * 1: LockNU), computed without any allocation call: no hipMalloc, hipFree,
 * pinned-memory or stream/event creation.  No counterpart in the reference;
 * it is for callers that cannot allocate (code that must not call hipMalloc:
 * tight serving loops, harnesses that forbid it), e.g. the per-block
 * transaction and receipt tries that build_mpt_2phase / build_mpt_olc build
 * on a fresh trie.  Every device buffer of the call comes from one fixed
 * device array of GMPT_NOALLOC_SCRATCH_MIB MiB per GPU, held until the process
 * ends; batches above GMPT_NOALLOC_MAX_KEYS keys, or needing more scratch
 * than the array holds, fail with an error (-3 / -2) and gmpt_last_error().
 * Host batches are staged into the array; device batches' values are read
 * in place.  Runs on the calling thread's current device and its null stream
 * and returns when the root is on the host: *size = 32, the raw root
 * encoding's length when it is under 32 bytes, or 0 for an empty batch.
 * Calls on one GPU share its array: calls from threads of one process take a
 * per-device lock and run one at a time; callers in other processes have
 * their own array.  The array keeps the last call's contents; each call
 * clears what it needs zeroed. */
#define GMPT_NOALLOC_MAX_KEYS 32768
#define GMPT_NOALLOC_SCRATCH_MIB 64
int gmpt_build_root_noalloc(const gmpt_batch *b, int mode, uint8_t out[32], int *size);

/* ---- Serving: continuous batching over HBM-resident tries ----------------
 * No reference counterpart (the reference builds one trie per call).  A
 * server holds num_tries resident tries split over `lanes` forests (one HIP
 * stream and one worker thread each; trie t lives in lane t % lanes).  A
 * request is one host batch for one trie; the server copies it at submit.
 * Each lane repeatedly inserts every queued request it can admit (at most one
 * per trie, so a trie's requests complete in submit order; total keys up to
 * max_batch_keys; node-pool bytes within mem_budget / lanes) as one
 * multi-trie batch (mode 0: PhaseNU, 1: LockNU), hashes once, and completes
 * each request with its trie's post-state root.  mem_budget = 0 takes 80% of
 * the device's free memory.  Request states: 0 queued, 1 running, 2 done,
 * -1 failed (gmpt_last_error() after the poll says why). */
typedef struct gmpt_server gmpt_server;
typedef struct gmpt_server_stats {
  uint64_t submitted, completed, failed;
  uint64_t batches, batched_requests, batched_keys;
  uint64_t deferred_mem;  /* admissions postponed by the memory budget */
  double busy_s;          /* summed lane time inside insert + hash */
} gmpt_server_stats;
int gmpt_server_create(int device, uint32_t num_tries, uint32_t lanes, uint64_t mem_budget, uint32_t max_batch_keys,
                       int mode, gmpt_server **out);
void gmpt_server_destroy(gmpt_server *s);
int gmpt_server_submit(gmpt_server *s, uint32_t trie, const gmpt_batch *b, uint64_t *req_id);
int gmpt_server_poll(gmpt_server *s, uint64_t req_id, int *state, uint8_t root[32], int *root_size,
                     double *latency_ms);
int gmpt_server_wait(gmpt_server *s, uint64_t req_id, int *state, uint8_t root[32], int *root_size,
                     double *latency_ms);
int gmpt_server_drain(gmpt_server *s);
int gmpt_server_get_stats(gmpt_server *s, gmpt_server_stats *out);
/* Latencies (ms, submit to root available) of completed requests, in request order. */
int gmpt_server_latencies(gmpt_server *s, double *out, uint64_t cap, uint64_t *count);

#ifdef __cplusplus
}
#endif

#endif /* GMPT_AMD_H */
