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
  double last_hash_ms;            /* device time of the last hash's Keccak kernels */
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
/* PhaseHC: re-hash every dirty node bottom-up and refresh each trie's root.
 * Replaces: MPT::hash_onepass_v2 (gpu_mpt.cuh:131, 1162-1191). */
int gmpt_hash_onepass(gmpt_forest *f, void *targets, int n_targets);
/* Root hash of one trie: *size = 32, or the raw root encoding length when it
 * is shorter than 32 bytes, or 0 for an empty trie.
 * Replaces: MPT::get_root_hash (gpu_mpt.cuh:151-152, 1217-1247). */
int gmpt_get_root_hash(gmpt_forest *f, uint32_t trie, uint8_t out[32], int *size);
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
int gmpt_get_stats(gmpt_forest *f, gmpt_stats *out);

/* Device helpers so callers can keep batches resident in HBM. */
int gmpt_device_count(int *count);
int gmpt_device_malloc(void **p, size_t bytes);
int gmpt_device_free(void *p);
int gmpt_memcpy_htod(void *dst, const void *src, size_t bytes);
int gmpt_memcpy_dtoh(void *dst, const void *src, size_t bytes);
/* Device-time stopwatch on the forest's stream. */
int gmpt_timer_start(gmpt_forest *f);
int gmpt_timer_stop(gmpt_forest *f, float *ms);
/* Keccak-256 of n host messages on the device (test hook for the sponge).
 * Replaces: GPUHashMultiThread batch hashing (batch_mode_hash.cuh:89-115). */
int gmpt_keccak256_batch(const uint8_t *data, const uint64_t *offsets, const uint32_t *lengths, uint32_t n,
                         uint8_t *out);

#ifdef __cplusplus
}
#endif

#endif /* GMPT_AMD_H */
