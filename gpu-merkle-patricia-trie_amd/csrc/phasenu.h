// phasenu.h -- PhaseNU: sort-based batch insert (host launch interface).
#pragma once
#include "common.h"

namespace gmpt {

// Where a key's descent into the existing trie stopped.
enum : uint8_t {
  STOP_ROOT_EMPTY = 0,   // the trie is empty: the key starts a new root
  STOP_BRANCH_SLOT = 1,  // an empty slot of an existing branch
  STOP_SHORT_SPLIT = 2,  // the key diverges inside an existing short node
  STOP_LEAF_UPDATE = 3,  // the key exists (leaf short node): replace its value
  STOP_VALUE_UPDATE = 4, // the key exists (value in a branch's slot 16)
};

struct Stop {
  ref_t stop;        // short node (split / leaf update), or branch (value update)
  ref_t attach;      // branch whose slot gets the new subtree, REF_NULL for the trie root
  uint16_t base;     // first nibble the new subtree covers
  uint16_t div;      // split: absolute nibble where the key leaves the short node
  uint8_t kind;
  uint8_t slot;
  uint8_t less;      // split: key orders before the short node's subtree
  uint8_t pad;
};

struct GroupInfo {
  ref_t attach;
  ref_t split;       // short node being split (REF_NULL if none)
  uint32_t trie;
  uint16_t base;
  uint8_t slot;
  uint8_t has_split;
};

// Device pointers of one PhaseNU call.  All arrays are sized by the host from
// n (the batch size); kernels read the actual counts (m, M) from device words.
struct NuArgs {
  Pools P;
  // batch (device)
  const uint8_t *kh;
  const int32_t *kidx;
  const int64_t *vidx;
  const uint64_t *hps;   // nullable
  const int32_t *tid;    // nullable
  uint32_t n;
  uint32_t ntries;
  uint64_t kh_bytes;     // bytes in kh (index pairs are checked against it)
  uint64_t v_bytes;      // bytes of the batch's values
  uint64_t kbase;        // this batch's offset in the key arena
  uint64_t vbase;        // device address of this batch's value bytes
  // per input key
  uint32_t *klen;
  uint64_t *kpoff;       // per block of 256 keys: its slot bytes, then (k_keys_reduce) its first slot's offset
  uint64_t *kref;        // kpoff | nibbles << 48: one load gives a key's place and length
  uint64_t key_cap;      // key arena bytes allocated (bounds the speculative packing)
  uint32_t *leaf_pk;     // per element: the new leaf's hashing bin (hc_leaf.h), or kLeafPackedNone
  uint32_t *bdesc;       // per new branch (by rank): depth | children << 16 | kBranchRecChase (hc_leaf.h)
  uint64_t *leaf_meta;   // per element: the new leaf's metadata words (two)
  // sort
  uint64_t *sk0, *sk1;
  uint32_t *sv0, *sv1;
  uint32_t *tk0, *tk1;
  uint32_t *scal;        // device scalars, see SC_*
  // sorted adjacent info
  uint8_t *lcp;          // lcp with previous sorted key (absolute nibbles), valid if !newtrie
  uint8_t *flags;        // bit0 dup-of-previous, bit1 new trie
  uint32_t *run;         // run id (inclusive scan result)
  // unique keys (m)
  uint32_t *u_idx, *u_trie;
  uint8_t *u_lcp, *u_first;  // lcp with previous unique key; first key of its trie
  Stop *stop;
  uint32_t *grp;         // group id of each unique key
  GroupInfo *ginfo;
  uint32_t *ecount;      // E entries emitted per unique key, scanned to positions
  uint32_t *merge;       // per run: run whose group it joins (split groups around a subtree)
  uint32_t *u_idx2, *u_trie2, *perm0, *perm1, *gkey0, *gkey1;
  // each unique key's kref and sort window, gathered in unique order so the
  // build reads them in sequence (u_win holds the 16 nibbles from W only when
  // u_win_ok: the last sort's keys were the plain windows)
  uint64_t *skref;       // per sorted position: kref of that key (null: not gathered)
  uint64_t *u_kref, *u_kref2, *u_win, *u_win2;
  uint32_t u_win_ok;
  uint8_t *u_lcp2;
  Stop *stop2;
  // E elements (M <= 2m)
  uint32_t *e_src;       // bit31: virtual (short node idx) else unique index
  uint32_t *e_grp;
  uint32_t e_ident;      // 1: a build into one empty trie, where element e is unique key e of group 0
                         // (e_src and e_grp are then not written)
  uint8_t *st;           // minimum tree over the split points (mt_* below); level 0 = split-point lcp+1
  size_t st_stride;      // bytes of level 0 (a multiple of 256)
  int st_levels;
  uint32_t *repof;       // per split point: representative split point (or ~0u)
  uint32_t *bid;         // per split point: branch rank (reps only)
  uint32_t *bcnt;        // per block of 256 split points: representatives (k_split_reps)
  uint64_t *plohi;       // per representative split point: its interval's bounding split points
                         // (nearest to the left and right with a smaller value) lo | hi << 32, -1 / nsp past the ends
  uint32_t *xp, *xm;     // links whose parent another block initialises: parent representative, child
  uint8_t *xs;           // and slot (512 entries per block of 256 split points)
  uint32_t *xcnt;        // per block: its cross links
  // LockNU
  uint64_t seq_base;     // batch order stamp of key 0
  uint32_t base_branch, base_short, base_value;  // pre-assigned node slots
};

// Minimum tree geometry: level 0 holds `stride` bytes (one per split
// point); level k holds one byte per 16 of level k - 1, padded to 16 bytes,
// right after it; the top level fits one 16-byte block.
__host__ __device__ inline size_t mt_level_len(size_t stride, int k) {
  size_t len = stride;
  for (int q = 0; q < k; q++) len = ((len / 16 + 15) / 16) * 16;
  return len;
}
__host__ __device__ inline int mt_levels(size_t stride) {
  int k = 1;
  while (mt_level_len(stride, k - 1) > 16) k++;
  return k;
}
__host__ __device__ inline size_t mt_bytes(size_t stride) {
  size_t b = 0;
  for (int k = 0; k < mt_levels(stride); k++) b += mt_level_len(stride, k);
  return b;
}

// device scalar slots in NuArgs::scal
enum {
  SC_W = 0, SC_MAXLEN = 1, SC_M_UNIQ = 2, SC_NGROUP = 3, SC_NE = 4, SC_KTOT = 6 /* u64 */, SC_MERGE = 8,
  SC_NREP = 9, SC_DBASE = 10, SC_NREP_TOT = 11,
  SC_WOR = 12 /* u64: OR of the windows */, SC_WAND = 14 /* u64: AND of the windows */,
  SC_UNIFORM = 16 /* 1: every key has key 0's length (packed by the validation pass) */,
  SC_NFRESH = 17 /* LockNU: keys the insert creates */,
  SC_DUP = 18 /* nonzero: the sorted batch holds equal keys (set by k_adjacent) */, SC_COUNT = 20
};

// Zeroes the call's device scalars (SC_W starts at all ones for its minimum).
hipError_t nu_init_scalars(const NuArgs &a, hipStream_t s);
hipError_t nu_prepare_keys(const NuArgs &a, void *temp, hipStream_t s);
// Sorts the batch: by the 16-nibble windows (full = 0; first by their top bits
// with a fix-up of equal runs, or by all 64 bits when win_all_bits) or by the
// whole keys (full = number of 16-nibble words), then by trie.  *cur_io (if
// given) is the ping-pong buffer holding the current order, in and out.
// With win_all_bits and a mask of the window bits that vary across the batch
// (vmask, from nu_window_mask), the windows are first compressed to those
// bits (an order-preserving bit gather to the top) and sorted by them only;
// *sk_mask receives the mask the returned keys are compressed with (0: raw
// windows), for nu_adjacent.
hipError_t nu_sort(NuArgs &a, int full, uint32_t trie_bits, void *temp, hipStream_t s, uint64_t **sk,
                   uint32_t **sv, bool win_all_bits = false, int *cur_io = nullptr, uint64_t vmask = 0,
                   uint64_t *sk_mask = nullptr);
// OR and AND of the n windows sk[0 .. n) into scal[SC_WOR], scal[SC_WAND]
// (per-block partials in a.ecount, which needs 4 words per 256 keys).
hipError_t nu_window_mask(const NuArgs &a, const uint64_t *sk, hipStream_t s);
hipError_t nu_adjacent(const NuArgs &a, const uint64_t *sk, const uint32_t *sv, int full, hipStream_t s,
                       uint64_t sk_mask = 0);
hipError_t nu_unique(const NuArgs &a, const uint64_t *sk, const uint32_t *sv, void *temp, hipStream_t s);
// Before a repeated sort pass (an eager fallback): clears SC_DUP, which the
// pass's k_adjacent sets again.
hipError_t nu_reset_dup(const NuArgs &a, hipStream_t s);
// Descends every unique key and groups them by stop point; sets SC_MERGE when
// a split group is interrupted by keys below the split node (nu_regroup).
// empty: every trie of the forest is empty (the host knows it), so the keys
// need no descent, the groups are the tries and the elements are emitted
// here too (nu_emit then does nothing, nu_build skips the update kernel).
hipError_t nu_descend_and_group(const NuArgs &a, void *temp, hipStream_t s, bool empty = false);
// Makes every group contiguous (stable sort of the unique keys by group),
// swapping the unique-key arrays of `a`, and regroups.
hipError_t nu_regroup(NuArgs &a, uint32_t m, uint32_t nruns, void *temp, hipStream_t s);
// Emits the element sequence E of every group.
hipError_t nu_emit(const NuArgs &a, void *temp, hipStream_t s, bool empty = false);
// Builds every group's subtree; node and dirty-list slots are fixed by
// position from base_branch / base_short / base_value (see phasenu.hip).
hipError_t nu_build(const NuArgs &a, void *temp, hipStream_t s, bool empty = false);

}  // namespace gmpt
