// phasehc.h -- PhaseHC: bottom-up one-pass hashing (host launch interface).
#pragma once
#include "common.h"

namespace gmpt {

constexpr uint32_t kHcClasses = 8;     // Keccak block-count classes per level
constexpr uint32_t kHcMaxDepth = 255;  // levels: 0 = leaves, 1 + (255 - depth) otherwise
constexpr uint32_t kHcLevels = kHcMaxDepth + 2;
constexpr uint32_t kHcBins = kHcLevels * kHcClasses;
// Levels with at most this many nodes spread each Keccak state over 25 lanes.
constexpr uint32_t kHcSpreadMaxDefault = 1024;
// Extensions per inner level that get their own (leaf word builder) kernel.
constexpr uint32_t kExtDirectMin = 16384;
// Spread-layout levels of at least this many nodes put two states in a wave.
constexpr uint32_t kSpread2Min = 2048;

// Classification results, read back by the host in one copy.
struct HcMeta {
  uint32_t hist[kHcBins];   // nodes per (level, class) bin
  uint32_t slot_lo;         // first sorted node with a scratch slot (past the level-0 leaves of hist[0 .. 8)
                            // when they are hashed without one); slot_lo, nvalid is the slot scan's range
  uint32_t nvalid;          // binned nodes (the dirty list minus null holes)
  uint64_t scratch_bytes;   // sum of the nodes' scratch slot sizes
};

size_t hc_temp_bytes(uint32_t n);
// Bins dirty nodes by (level, block class): out = the list ordered by bin,
// slot[i] = byte offset of sorted node i's message slot in the scratch buffer
// (for slot_lo <= i < nvalid; other entries are not written);
// packed (n entries) is workspace.  With lmeta (2n words, lmeta_tmp 2n words
// of workspace) the leaves whose key remainder has at most 64 nibbles and
// whose value is under 16,383 bytes -- level 0, the first hist[0 .. 8)
// sorted entries -- get no scratch slot but a metadata record lmeta[2i],
// lmeta[2i + 1] for hc_hash_leaves_direct; other leaves are binned at
// level 1 (no inner node can be there).
// Bins and leaf metadata recorded by a build for dirty-list entries
// [lo, lo + n): pk[k] = bin | blocks << 12 or kLeafPackedNone, meta[2k..] as
// leaf_meta_words.  Valid only while nothing has changed those nodes since.
struct LeafRecords {
  const uint32_t *pk = nullptr;
  const uint64_t *meta = nullptr;
  uint32_t lo = 0, n = 0;
  // the build's new branches, dirty-list entries [blo, blo + bn): bdesc[k] as
  // NuArgs::bdesc (depth, child count, whether to derive the bin from the node)
  const uint32_t *bdesc = nullptr;
  uint32_t blo = 0, bn = 0;
};
hipError_t hc_classify(const Pools &P, const ref_t *list, uint32_t n, HcMeta *meta, uint32_t *cursor, ref_t *out,
                       uint64_t *slot, uint32_t *packed, void *temp, hipStream_t s, uint64_t *lmeta_tmp = nullptr,
                       uint64_t *lmeta = nullptr, LeafRecords rec = LeafRecords());
// A second stream (default priority) and two events for work that runs
// beside the main stream's kernels within one hash pass.
struct SideStream {
  hipStream_t stream;
  hipEvent_t fork, join;
};
// Encodes and hashes list[0..n) (one level: no node depends on another);
// slot and enc_len are the level's slices of the classification arrays;
// leaves = the leaf levels (0 and 1: only leaf short nodes).  Of an inner
// level, the last n_ext entries are its extension class, hashed by the leaf
// word builder when the level is above spread_max.  Events, when given, are
// recorded around the permutation kernel.  With cls (the level's node counts
// per class, from the classification's bins), an inner level above
// spread_max of 1- and 2-block branches and extensions (all but a sixteenth
// of its nodes) is encoded and hashed with its messages in LDS in one launch
// (k_hc_inner_wave); other levels, or all without cls, by a 16-lane encoding
// kernel, an HBM scratch message and a permutation kernel.  The LDS kernel's
// first blocks hash the level's few nodes taller than its rows, two per
// block, beside the bulk.
hipError_t hc_hash_range(const Pools &P, const ref_t *list, const uint64_t *slot, uint32_t n, uint64_t *scratch,
                         uint32_t *enc_len, bool leaves, uint32_t spread_max, hipStream_t s,
                         hipEvent_t perm_begin = nullptr, hipEvent_t perm_end = nullptr, uint32_t n_ext = 0,
                         const uint32_t *cls = nullptr);
// Encodes and hashes the classification's level-0 entries (leaves with
// metadata, see hc_classify; their count is the sum of meta->hist[0 .. 8))
// in one kernel; the count is read on the device, n_upper (>= it) sizes the
// launch, so the launch need not wait for the bins to reach the host.
hipError_t hc_hash_leaves_direct(const Pools &P, const ref_t *list, const uint64_t *lmeta, const HcMeta *meta,
                                 uint32_t n_upper, uint32_t *enc_len, hipStream_t s);
// Hashes the leaf children of the inner nodes list[0 .. n) (the deepest
// levels, taken off the critical path onto a second stream), every one a
// leaf the word builder takes.
hipError_t hc_tail_leaves(const Pools &P, const ref_t *list, uint32_t n, hipStream_t s);
// The deepest inner levels a pass may hash beside its leaf-level launch: at
// most this many nodes in all, none of them in a level above kTailLevelMax.
constexpr uint32_t kTailMax = 8192;
constexpr uint32_t kTailLevelMax = 4096;
// The pass's permutations and encoding bytes (from the enc_len array of the
// meta->nvalid hashed nodes), and those of the first n_leaf entries (the
// leaf level), as kHcCountParts per-block partial sums (4 words each) in
// parts; hc_roots adds them to the forest counters.
constexpr uint32_t kHcCountParts = 256;
hipError_t hc_count(const uint32_t *enc_len, const HcMeta *meta, uint32_t n, uint32_t n_leaf,
                    unsigned long long *parts, hipStream_t s);
// Index of the element count among an insert's device scalars (PhaseNU's SC_NE).
constexpr uint32_t SC_NE_INDEX = 4;
// Replay check of a captured insert + hash pass: marks the pass bad
// (Counters::graph_bad, which the hash kernels test) unless the batch sorted
// without fallback, raised no error and gave n_dirty dirty nodes, ne elements
// (scal[SC_NE_INDEX]) and the bins plan_hist (kHcBins counts, device).
hipError_t hc_graph_check(const Pools &P, const HcMeta *meta, const uint32_t *plan_hist, uint32_t n_dirty, uint32_t ne,
                          const uint32_t *scal, hipStream_t s);
// Zeroes the forest's counters and every trie header (empty tries).
hipError_t hc_clear(Counters *cnt, TrieHdr *tries, uint32_t ntries, hipStream_t s);
// Copies every trie's root hash into its header and empties the dirty list;
// with parts (from hc_count), adds the pass's work counts to the counters.
hipError_t hc_roots(const Pools &P, uint32_t ntries, const unsigned long long *parts, hipStream_t s);
// Node export: encoding lengths, then encodings at offs and 32-byte hash slots.
hipError_t hc_enc_len(const Pools &P, const ref_t *refs, uint32_t m, uint64_t *len, hipStream_t s);
hipError_t hc_enc_write(const Pools &P, const ref_t *refs, uint32_t m, const uint64_t *offs, uint8_t *encs,
                        uint8_t *hashes, hipStream_t s);
// Keccak-256 of n messages (test hook for the sponge): scratch holds
// keccak_scratch_bytes(len, n) bytes, slot n entries, temp scan workspace.
// spread = 1 runs the 25-lane layout, 0 one state per lane.
uint64_t keccak_scratch_bytes(const uint32_t *len, uint32_t n);
hipError_t keccak_batch(const uint8_t *data, const uint64_t *off, const uint32_t *len, uint32_t n, uint8_t *out,
                        uint64_t *slot, uint64_t *scratch, void *temp, int spread, hipStream_t s);

}  // namespace gmpt
