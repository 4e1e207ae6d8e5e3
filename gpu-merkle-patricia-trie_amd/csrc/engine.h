// engine.h -- host runtime: a forest of tries resident in HBM.
//
// One Forest owns the node pools, the key arena, value chunks and the tries'
// headers on one device, and runs every operation on its own HIP stream.  A
// single trie (the reference's GpuMPT::Compress::MPT, gpu_mpt.cuh:18-192) is a
// forest of one; batched and served workloads use one forest for many tries.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "common.h"
#include "phasehc.h"
#include "phasenu.h"

namespace gmpt {

class HostStager;
struct StageSeg;

struct BatchDesc {
  const uint8_t *keys_hexs = nullptr;
  const int32_t *keys_indexs = nullptr;    // inclusive [start, end] pairs
  const uint8_t *values_bytes = nullptr;
  const int64_t *values_indexs = nullptr;  // inclusive [start, end] pairs
  const uint64_t *values_hps = nullptr;    // caller's host pointers (nullable)
  const int32_t *trie_ids = nullptr;       // nullable: all keys go to trie 0
  uint32_t n = 0;
  uint64_t keys_bytes = 0;                 // totals (computed from the indexs for host batches)
  uint64_t values_total = 0;
  uint64_t key_slots = 0;                  // key arena bytes of the packed keys (set by stage_batch)
  bool on_device = false;                  // all pointers above are device pointers
  bool adopt_values = false;               // device batch: nodes point into values_bytes (caller keeps it alive)
};

struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  bool host = false;  // pinned host memory mapped into the device's address space
};

// A fixed device region a Forest can take all its device memory from, with
// no allocation call: buffers are carved in order and never returned (the
// forest lives for one call).  take() gives nullptr once the region is spent.
struct Arena {
  uint8_t *base = nullptr;
  size_t cap = 0;
  size_t used = 0;
  void *take(size_t bytes) {
    size_t off = (used + 255) & ~(size_t)255;
    if (off > cap || bytes > cap - off) return nullptr;
    used = off + bytes;
    return base + off;
  }
};

// Host memory a Forest over an Arena copies its counters and hashing bins
// into (pageable; the caller owns it, so the forest allocates none).
struct HostMirror {
  Counters cnt;
  uint32_t scal[16];
  HcMeta meta;
};

struct OpStats {
  double last_hash_ms = 0;    // device time of the hash kernels of the last hash()
  uint32_t hash_launches = 0;
  uint32_t hashed_nodes = 0;
  uint32_t restarts = 0;      // LockNU optimistic restarts (last op)
  uint32_t lock_fails = 0;    // LockNU failed latch upgrades (last op)
  uint32_t fallback_sorts = 0;
  uint32_t regroups = 0;       // PhaseNU group regroupings (split groups around a subtree)
  uint64_t perms = 0;          // Keccak-f permutations of the last hash
  uint64_t enc_bytes = 0;      // node encoding bytes of the last hash
  double leaf_perm_ms = 0;     // device time of the leaf level's permutation kernel
  uint64_t leaf_perms = 0;     // permutations of the leaf level
  uint64_t leaf_bytes = 0;     // encoding bytes of the leaf level
  uint32_t leaf_nodes = 0;
  uint32_t ext_direct_nodes = 0;  // extensions hashed by the leaf word builder
};

class Forest {
 public:
  Forest(int device, uint32_t ntries);
  // A forest that allocates nothing: its device memory comes from `arena`,
  // its host mirrors from `host`, and it runs on `stream` (not owned).  It
  // keeps no timing events, so its stats carry no kernel times.
  Forest(int device, uint32_t ntries, Arena *arena, HostMirror *host, hipStream_t stream);
  ~Forest();
  bool ok() const { return init_ok_; }

  // PhaseNU (sort-based) and LockNU (latch-coupled) batch inserts.  On success
  // *targets / *n_targets name the dirty list to pass to hash().
  int puts_2phase(const BatchDesc &b, void **targets, int *n_targets);
  int puts_latching(const BatchDesc &b, void **targets, int *n_targets);
  // PhaseHC over the dirty list, then every trie's root hash is refreshed.
  int hash(void *targets, int n_targets);
  // Mixed batch: rw_flags[i] = 0 reads key i, 1 writes it (host or device
  // array, as the batch).  Reads see the tries as they were before the batch
  // (all reads are ordered before all writes); their results come back in
  // batch order of the reads (*read_num of them).  Writes then run as one
  // PhaseNU (mode 0) or LockNU (mode 1) insert.
  int puts_with_read(const BatchDesc &b, const uint8_t *rw_flags, int mode, uint32_t *read_num,
                     uint64_t *read_hps, int32_t *read_sizes, void **targets, int *n_targets);
  int root_hash(uint32_t trie, uint8_t *out, int *size);
  // Root hashes of tries [first, first + count) in one copy: 32-byte slots
  // (zero padded) and sizes.
  int root_hashes(uint32_t first, uint32_t count, uint8_t *out, int32_t *sizes);
  // get: value size (-1 absent) and the caller's host pointer per key.
  int gets(const BatchDesc &keys, uint64_t *hps, int32_t *sizes);
  // get with value bytes copied out (offsets per key into out).
  int get_values(const BatchDesc &keys, uint8_t *out, int64_t out_cap, int64_t *offsets, int32_t *sizes);
  // Every node on the paths of the given keys: 32-byte hash slots and RLP
  // encodings with inclusive [start, end] index pairs (reference nodeset).
  int export_nodes(const BatchDesc &keys, uint8_t **hashes, uint8_t **encs, uint64_t **enc_indexs,
                   uint64_t *num);
  int reset_trie(uint32_t trie);
  // Empties every trie and recycles all pool memory (no reallocation).
  int clear();
  int synchronize();
  // Tunables by name ("hc_spread_max": levels with at most this many dirty
  // nodes use the spread Keccak layout).
  int set_option(const std::string &name, int64_t value);
  int get_option(const std::string &name, int64_t *value) const;
  // Times PhaseHC on `sample` (inserted into a scratch forest on the same
  // device) for each candidate spread threshold and keeps the fastest.
  int autotune(const BatchDesc &sample, int mode, int64_t *chosen, double *best_ms);
  // Bulk rebuild of a device batch (values adopted) into the emptied forest
  // -- clear, PhaseNU, PhaseHC -- as one replay of a captured HIP graph.  The
  // first call for a batch (its buffers and sizes) runs eagerly and captures
  // the pass; later calls replay it, and a replay whose batch turns out unlike
  // the captured one (sort fallback, other hash levels) is redone eagerly and
  // captured again.  *replayed tells which ran.
  int rebuild_graph(const BatchDesc &b, bool *replayed);
  uint32_t ntries() const { return ntries_; }
  uint64_t pool_bytes_used() const;
  uint64_t pool_bytes_reserved() const;
  const OpStats &stats() const { return stats_; }
  hipStream_t stream() const { return stream_; }
  int device() const { return device_; }
  Pools pools() const;

 private:
  int set_device();
  int ensure_pools(uint64_t nb, uint64_t ns, uint64_t nv, uint64_t key_bytes);
  int grow(DevBuf &b, size_t need, size_t keep);
  int ensure_ws(DevBuf &b, size_t bytes);
  int arena_exhausted();
  int hash_levels(const Pools &P, const uint32_t *hist, bool direct_leaves, bool captured, uint32_t *n_leaf_out);
  // rebuild_graph's parts (graph.hip)
  int rebuild_eager(const BatchDesc &b);
  int capture_rebuild(const BatchDesc &b);
  int replay_graph();
  bool graph_ok();
  void drop_graph();
  bool is_pool(const DevBuf &b) const;
  void release(DevBuf &b);
  // Reads the counters back (raising a kernel's error); with mark, the host
  // waits for that event, recorded after the read, instead of the stream.
  int sync_counters(hipEvent_t mark = nullptr);
  bool grow_key_arena_after_error();
  // With defer, a host batch large enough for the pinned stager leaves its
  // values out and describes them in *defer (bytes 0 when nothing is left):
  // the caller stages them later, before anything reads value bytes.
  int stage_batch(const BatchDesc &b, BatchDesc *dev, uint64_t *vbase, bool copy_values, StageSeg *defer = nullptr);
  int carve_nu(NuArgs &a, const BatchDesc &d, uint64_t vbase, bool has_tid, void **temp);
  int sort_unique(NuArgs &a, uint32_t trie_bits, void *temp, int pass, int *cur);
  int next_pass(int pass);
  uint8_t *value_space(uint64_t bytes);

  int device_ = 0;
  uint32_t ntries_ = 0;
  bool init_ok_ = false;
  Arena *arena_ = nullptr;     // set: every device buffer is carved from it
  bool all_empty_ = true;      // no insert since creation or clear(): every trie is empty
  hipStream_t stream_ = nullptr;
  DevBuf branch_, shrt_, value_, keys_, tries_, dirty_;
  Counters *d_cnt_ = nullptr;
  Counters *h_cnt_ = nullptr;  // pinned
  uint32_t *h_scal_ = nullptr; // pinned
  Counters cnt_{};             // host mirror after the last sync
  uint64_t key_used_ = 0;
  uint64_t lock_seq_ = 1;    // LockNU order stamps (PhaseNU values carry 0)
  std::vector<DevBuf> vchunks_;
  size_t vchunk_used_ = 0;
  DevBuf st_keys_, st_kidx_, st_vidx_, st_hps_, st_tid_;
  DevBuf ws_, ws2_, rw_;
  HostStager *stager_ = nullptr; // pinned chunked staging of large host batches (made on first use)
  DevBuf hc_sorted_, hc_meta_, hc_cursor_, hc_slot_, hc_enc_, hc_scratch_, hc_lmeta_, hc_lmeta_tmp_;
  void *h_meta_ = nullptr;       // pinned HcMeta mirror
  bool root0_ok_ = false;        // cnt_.root0 is trie 0's current root (set by hash, cleared by every other call)
  uint32_t hc_spread_max_ = 0;   // levels up to this size use the spread permutation
  bool hc_fuse_leaves_ = true;   // level-0 leaves: one encode+hash kernel (k_hc_leaf_direct)
  bool hc_fuse_inner_ = true;    // large inner levels: encode + hash with the messages in LDS (k_hc_inner_wave)
  // "pool_memory" 1: the trie state (node pools, key and value arenas, dirty
  // list, trie headers) grows in pinned host memory the kernels reach over
  // the host link, so a forest can hold more than the GPU's memory
  bool pool_host_ = false;
  LeafRecords rec_;              // bins the last build recorded for its new leaves (in ws_)
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;  // around a hash pass (after the bins' copy; after the counters' read)
  hipEvent_t ev2_ = nullptr, ev3_ = nullptr;  // around the leaf level's permutation kernel
  hipEvent_t ev_meta_ = nullptr;               // the classification bins have reached the host
  hipEvent_t meta_mark_ = nullptr;             // the event hash() recorded after the bins' copy (ev0_ or ev_meta_)
  hipEvent_t tail_join_ = nullptr;             // the second stream has hashed the pass's deepest levels
  SideStream side_{};                          // second stream for work beside a hash level's main kernel
  // the captured rebuild pass and the batch it was captured for
  struct GraphPlan {
    bool valid = false;
    BatchDesc batch;               // buffers and sizes the graph's launches use
    uint32_t n_dirty = 0, ne = 0;  // dirty nodes and elements of the captured pass
    uint64_t key_used = 0;         // key arena bytes after it
    uint32_t hist[kHcBins];        // its hash bins
  };
  GraphPlan plan_;
  DevBuf plan_hist_;                 // plan_.hist on the device, for the replay check
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t gexec_ = nullptr;
  OpStats stats_;
};

void set_error(const std::string &e);
const char *last_error();

}  // namespace gmpt
