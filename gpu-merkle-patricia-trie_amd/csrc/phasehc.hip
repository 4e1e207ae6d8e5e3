// phasehc.hip -- PhaseHC: bottom-up one-pass hashing of the dirty nodes.
//
// Reference: hash_onepass_v2 (mpt-with-compress/include/mpt/gpu_mpt.cuh:1162-1191)
// counts children per node (mark phase, gpu_mpt_kernels.cuh:2696-2728) and
// lets the last-arriving warp encode and hash each node (update phase,
// :2885-3241).  Here every dirty node carries its nibble depth, so the hash
// order is fixed without counters: all leaf short nodes first (they depend on
// nothing), then non-leaf nodes by decreasing depth (a child is always deeper
// than its parent).  Nodes are binned by (level, estimated Keccak blocks), so
// the lanes of a wave run similar numbers of permutations.
//
// Per level two kernels run:
//  - encode: one thread per node writes the node's RLP encoding, Keccak
//    padding included, into its slot of an HBM scratch buffer (slot sizes are
//    the binning's upper bound).  An encoding shorter than 32 bytes becomes
//    the node's "hash" directly.  The kernel holds few registers, so many
//    waves overlap the pointer chasing to children and values.
//  - permute: Keccak-f[1600] over the slots.  Large levels keep one 25-word
//    state per GPU lane (VGPRs, no communication); small levels are latency
//    bound, so there one state is spread over 25 lanes of a 32-lane group
//    and theta, pi and chi exchange words with cross-lane permutes.
//
// Node encoding follows the reference's encoders exactly
// (FullNode::encode node.cuh:414-452, ShortNode::encode node.cuh:526-559,
// rlp.cuh:118-155): every child is written with write_bytes, i.e. a child
// whose encoding is shorter than 32 bytes is string-wrapped (not inlined),
// a value node contributes its raw bytes, and a node whose own encoding is
// shorter than 32 bytes keeps the raw encoding instead of a digest.
#include "phasehc.h"

#include <algorithm>

#include "hc_leaf.h"
#include "keccak.h"
#include "prims.h"

namespace gmpt {

namespace {

constexpr int kB = 256;
constexpr uint32_t kHcTile = 2048;  // dirty nodes per binning block

__device__ __forceinline__ uint32_t grid_tid() { return blockIdx.x * blockDim.x + threadIdx.x; }
inline unsigned nblk(size_t n) { return (unsigned)((n + kB - 1) / kB); }


// header bytes packed little-endian into a u64 (first byte lowest)
__device__ __forceinline__ uint64_t rlp_hdr(uint32_t L, uint32_t small, uint32_t large, uint32_t *n) {
  if (L < 56) {
    *n = 1;
    return small + L;
  }
  uint32_t nb = be_bytes(L);
  uint64_t h = large + nb;
  for (uint32_t q = 0; q < nb; q++) h |= (uint64_t)((L >> (8 * (nb - 1 - q))) & 255u) << (8 * (q + 1));
  *n = nb + 1;
  return h;
}


// ---------------------------------------------------------- classification --

// Keccak blocks of a node's message, bounded from above before its children
// are hashed: a hashed child counts as a 32-byte digest, and an embedded
// child (encoding < 32 bytes) is never longer than that.
__device__ uint32_t estimate_blocks(const Pools &P, ref_t r, bool *leaf) {
  uint32_t t = ref_type(r), i = ref_idx(r);
  uint32_t len;
  *leaf = false;
  if (t == T_SHORT) {
    const Short &S = P.shrt[i];
    if (ref_type(S.child) == T_VALUE) {
      *leaf = true;
      return leaf_blocks((uint32_t)S.key_len - S.term, P.value[ref_idx(S.child)].size);
    }
    const uint32_t pay = rlp_item_len(((uint32_t)S.key_len - S.term) / 2 + 1, 0) + 33;
    len = rlp_hdr_len(pay) + pay;
  } else {
    const Branch &B = P.branch[i];
    uint32_t pay = 0;
#pragma unroll
    for (int s = 0; s < 17; s++) {
      ref_t c = B.child[s];
      if (c == REF_NULL)
        pay += 1;
      else if (ref_type(c) == T_VALUE)
        pay += rlp_item_len(P.value[ref_idx(c)].size, 0x80);
      else
        pay += 33;
    }
    len = rlp_hdr_len(pay) + pay;
  }
  return keccak_blocks(len);
}

// Inner nodes have at most 5 blocks, so an inner level's last class is free
// for the extensions k_hc_ext_direct hashes: a short node over a branch with
// a key of 1 .. 64 nibbles.
constexpr uint32_t kExtClass = kHcClasses - 1;
__device__ __forceinline__ bool is_direct_ext(const Short &S) {
  return !S.term && ref_type(S.child) == T_BRANCH && S.key_len >= 1 && S.key_len <= kLeafFastNibbles;
}

__device__ __forceinline__ uint32_t bin_of(const Pools &P, ref_t r, uint32_t *blocks) {
  bool leaf;
  uint32_t nb = estimate_blocks(P, r, &leaf);
  *blocks = nb;
  uint32_t cls = nb > kHcClasses ? kHcClasses - 1 : nb - 1;
  if (!leaf && ref_type(r) == T_SHORT && is_direct_ext(P.shrt[ref_idx(r)])) cls = kExtClass;
  uint32_t level;
  if (leaf) {
    level = 0;
  } else {
    uint32_t depth = ref_type(r) == T_SHORT ? P.shrt[ref_idx(r)].depth : P.branch[ref_idx(r)].depth;
    level = 1 + (kHcMaxDepth - min(depth, (uint32_t)kHcMaxDepth));
  }
  return level * kHcClasses + cls;
}

// Leaf metadata (leaf_meta_words) for the leaf-level kernel, recorded by the
// classification in sorted order so that kernel's loads do not chase
// list -> short -> value.
__device__ __forceinline__ void leaf_meta(const Pools &P, ref_t r, uint64_t *m) {
  const Short &S = P.shrt[ref_idx(r)];
  const Value &V = P.value[ref_idx(S.child)];
  leaf_meta_words(S.key_off, S.key_start, (uint32_t)S.key_len - S.term, S.term, V.size, V.addr, m);
}
// Level 0 holds the leaves k_hc_leaf_direct hashes (key remainder <= 64
// nibbles, value < 16,383 bytes); other leaves go to level 1, which no inner
// node uses (a node at nibble depth 255 cannot exist), and take the encode +
// scratch + permutation kernels.
constexpr uint32_t kLdsLeafBins = kHcClasses;  // the bins of level 0

__global__ __launch_bounds__(256) void k_hc_zero(uint32_t *p, uint32_t n, uint32_t *p2 = nullptr, uint32_t n2 = 0) {
  uint32_t i = grid_tid();
  if (i < n) p[i] = 0;
  if (i < n2) p2[i] = 0;
}

// Counting sort of the dirty list by bin.  The first pass computes each
// entry's bin and block estimate once (pointer chasing) and keeps them packed
// in `packed` (bin | blocks << 12, blocks saturated at 2^20 - 1); counts are
// aggregated per block in LDS so each block issues one global atomic per bin
// it touches (a global atomic per node serialises on the few hot bins).

// A leaf whose key remainder or value is beyond the word builder's reach
// (metadata word m0) needs the generic encoder and a scratch slot.
__device__ __forceinline__ bool leaf_too_long(uint64_t m0) {
  return ((m0 >> 41) & 255u) > kLeafFastNibbles || (m0 >> 50) > kLeafDirectMaxValue;
}

// Each thread takes kHcTile / 256 entries; their list and leaf-record loads
// are issued together before any is used (the block count is kept low for
// the global histogram, so each thread's loads must overlap).
constexpr int kHcPer = (int)(kHcTile / 256);
// `packed` entries: bin | blocks << 12 (saturated at kPackedNbMax), and
// kPackedRecorded for a leaf the last build recorded: its metadata is read
// from the build's records, not from lmeta_tmp.
constexpr uint32_t kPackedRecorded = 1u << 31;

__global__ __launch_bounds__(256) void k_hc_hist(Pools P, const ref_t *__restrict__ list, uint32_t n,
                                                 uint32_t *__restrict__ hist, uint32_t *__restrict__ packed,
                                                 uint64_t *__restrict__ lmeta_tmp, LeafRecords rec) {
  __shared__ uint32_t h[kHcBins];
  for (uint32_t b = threadIdx.x; b < kHcBins; b += blockDim.x) h[b] = 0;
  const uint32_t i0 = blockIdx.x * kHcTile + threadIdx.x;
  ref_t r[kHcPer];
  uint32_t rpk[kHcPer];
#pragma unroll
  for (int q = 0; q < kHcPer; q++) {
    const uint32_t i = i0 + 256u * q;
    const uint32_t ri = i - rec.lo;
    r[q] = i < n ? list[i] : REF_NULL;  // builds leave null holes
    // a leaf the last build created: its bin (level 1 when the word builder
    // cannot take it) and metadata were recorded then
    rpk[q] = i < n && rec.pk && ri < rec.n ? rec.pk[ri] : kLeafPackedNone;
  }
  // a branch the last build created: its bin from the depth and child count
  // recorded then (or kLeafPackedNone: read the node)
  uint32_t bpk[kHcPer];
#pragma unroll
  for (int q = 0; q < kHcPer; q++) {
    const uint32_t i = i0 + 256u * q;
    const uint32_t bi = i - rec.blo;
    bpk[q] = i < n && rec.bdesc && bi < rec.bn ? branch_packed(rec.bdesc[bi]) : kLeafPackedNone;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kHcPer; q++) {
    const uint32_t i = i0 + 256u * q;
    if (i >= n) break;
    uint32_t pk = 0xffffffffu;
    if (r[q] != REF_NULL) {
      uint32_t nb, bin, flag = 0;
      if (rpk[q] != kLeafPackedNone) {
        bin = rpk[q] & 0xfff;
        nb = (rpk[q] >> 12) & kPackedNbMax;
        flag = kPackedRecorded;
        // without the word builder every leaf is hashed at level 0
        if (!lmeta_tmp && bin >= kHcClasses) bin -= kHcClasses;
      } else if (bpk[q] != kLeafPackedNone) {
        bin = bpk[q] & 0xfff;
        nb = (bpk[q] >> 12) & kPackedNbMax;
      } else {
        bin = bin_of(P, r[q], &nb);
        if (lmeta_tmp && bin < kLdsLeafBins) {
          uint64_t *m = lmeta_tmp + 2 * (size_t)i;
          leaf_meta(P, r[q], m);
          if (leaf_too_long(m[0])) bin += kHcClasses;
        }
      }
      atomicAdd(&h[bin], 1u);
      pk = bin | (min(nb, kPackedNbMax) << 12) | flag;
    }
    packed[i] = pk;
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < kHcBins; b += blockDim.x)
    if (h[b]) atomicAdd(&hist[b], h[b]);
}

// Scatters each node to its bin and records its scratch slot size
// (136 bytes per estimated block) at the same sorted position.
__global__ __launch_bounds__(256) void k_hc_scatter(Pools P, const ref_t *__restrict__ list,
                                                    const uint32_t *__restrict__ packed, uint32_t n,
                                                    uint32_t *__restrict__ cursor, ref_t *__restrict__ out,
                                                    uint64_t *__restrict__ slot,
                                                    const uint64_t *__restrict__ lmeta_tmp,
                                                    uint64_t *__restrict__ lmeta, LeafRecords rec, HcMeta *meta) {
  __shared__ uint32_t cnt[kHcBins];
  __shared__ uint32_t base[kHcBins];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // the slot scan's range starts past the level-0 leaves hashed without a slot
    uint32_t lo = 0;
    if (lmeta)
      for (uint32_t c = 0; c < kLdsLeafBins; c++) lo += meta->hist[c];
    meta->slot_lo = lo;
  }
  for (uint32_t b = threadIdx.x; b < kHcBins; b += blockDim.x) cnt[b] = 0;
  __syncthreads();
  constexpr int kPer = kHcTile / 256;
  uint32_t pk[kPer], rank[kPer];
#pragma unroll
  for (int q = 0; q < kPer; q++) {
    uint32_t i = blockIdx.x * kHcTile + q * 256 + threadIdx.x;
    pk[q] = i < n ? packed[i] : 0xffffffffu;
    if (pk[q] != 0xffffffffu) rank[q] = atomicAdd(&cnt[pk[q] & 0xfff], 1u);
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < kHcBins; b += blockDim.x)
    if (cnt[b]) base[b] = atomicAdd(&cursor[b], cnt[b]);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kPer; q++)
    if (pk[q] != 0xffffffffu) {
      uint32_t i = blockIdx.x * kHcTile + q * 256 + threadIdx.x;
      uint32_t pos = base[pk[q] & 0xfff] + rank[q];
      ref_t r = list[i];
      uint32_t nb = (pk[q] >> 12) & kPackedNbMax;
      if (nb == kPackedNbMax) bin_of(P, r, &nb);  // saturated: recompute exactly
      out[pos] = r;
      if (lmeta && (pk[q] & 0xfff) < kLdsLeafBins) {  // hashed by k_hc_leaf_direct: no scratch slot
        const uint64_t *m = (pk[q] & kPackedRecorded) ? rec.meta + 2 * (size_t)(i - rec.lo) : lmeta_tmp + 2 * (size_t)i;
        *(uint4 *)(lmeta + 2 * (size_t)pos) = *(const uint4 *)m;
        nb = 0;
      }
      slot[pos] = 136ull * nb;
    }
}


// ------------------------------------------------------------------ encode --

// Unaligned 8-byte read as aligned loads plus a byte funnel shift.
__device__ __forceinline__ uint64_t load8(const uint8_t *p) {
  uintptr_t a = (uintptr_t)p;
  uint32_t sh = (uint32_t)(a & 7);
  const uint64_t *q = (const uint64_t *)(a - sh);
  uint64_t w0 = q[0];
  if (sh == 0) return w0;
  return (w0 >> (8 * sh)) | (q[1] << (64 - 8 * sh));
}
// The first c (1..7) bytes at p, upper bytes zero; reads only words holding them.
__device__ __forceinline__ uint64_t load_part(const uint8_t *p, uint32_t c) {
  uintptr_t a = (uintptr_t)p;
  uint32_t sh = (uint32_t)(a & 7);
  const uint64_t *q = (const uint64_t *)(a - sh);
  uint64_t v = q[0] >> (8 * sh);
  if (sh + c > 8) v |= q[1] << (64 - 8 * sh);
  return v & ((1ull << (8 * c)) - 1);
}

// Writes a byte stream as aligned little-endian 64-bit words.
struct WordWriter {
  uint64_t *out;
  uint64_t acc = 0;   // pending bytes
  uint32_t accn = 0;  // number of pending bytes (0..7)
  uint32_t nw = 0;    // words written
  __device__ __forceinline__ explicit WordWriter(uint64_t *o) : out(o) {}
  // appends the low n (1..8) bytes of v; v's bytes above n must be zero
  __device__ __forceinline__ void push(uint64_t v, uint32_t n) {
    acc |= v << (8 * accn);
    uint32_t tot = accn + n;
    if (tot >= 8) {
      out[nw++] = acc;
      uint32_t used = 8 - accn;
      acc = used >= 8 ? 0ull : (v >> (8 * used));
      accn = tot - 8;
    } else {
      accn = tot;
    }
  }
  __device__ __forceinline__ void inline_bytes(uint64_t b, uint32_t len) {  // len 1..8
    push(len < 8 ? b & ((1ull << (8 * len)) - 1) : b, len);
  }
  __device__ __forceinline__ void mem_bytes(const uint8_t *src, uint32_t len) {
    for (; len >= 8; len -= 8, src += 8) push(load8(src), 8);
    if (len) push(load_part(src, len), len);
  }
  // write_bytes (rlp.cuh:132-155) of a memory string
  __device__ __forceinline__ void rlp_mem(const uint8_t *src, uint32_t len) {
    if (len == 1 && src[0] <= 0x7f) {
      inline_bytes(src[0], 1);
      return;
    }
    uint32_t hn;
    uint64_t h = rlp_hdr(len, 0x80, 0xB7, &hn);
    inline_bytes(h, hn);
    mem_bytes(src, len);
  }
  // write_bytes of the hex-prefix form (utils.cuh:372-396) of nibbles
  // [s, s + cnt) of a packed key (high nibble first) with the leaf flag term
  __device__ __forceinline__ void rlp_compact(const uint8_t *key, uint32_t s, uint32_t cnt, uint32_t term) {
    uint32_t kc = cnt / 2 + 1;
    uint32_t flag = (term << 5) | ((cnt & 1) << 4) | ((cnt & 1) ? packed_nibble(key, s) : 0u);
    if (kc == 1) {  // a single byte <= 0x3f is written raw
      inline_bytes(flag, 1);
      return;
    }
    uint32_t hn;
    uint64_t h = rlp_hdr(kc, 0x80, 0xB7, &hn);
    inline_bytes(h, hn);
    inline_bytes(flag, 1);
    uint32_t first = s + (cnt & 1);
    uint32_t c = kc - 1;
    if ((first & 1) == 0) {
      mem_bytes(key + first / 2, c);
      return;
    }
    // pairs straddle bytes: out byte i = low nibble of b[i], high nibble of b[i+1]
    const uint8_t *sp = key + first / 2;
    for (; c > 0; sp += 8) {
      uint64_t x = load8(sp), y = load8(sp + 1);
      uint64_t v = ((x << 4) & 0xF0F0F0F0F0F0F0F0ull) | ((y >> 4) & 0x0F0F0F0F0F0F0F0Full);
      uint32_t t = c < 8 ? c : 8;
      if (t < 8) v &= (1ull << (8 * t)) - 1;
      push(v, t);
      c -= t;
    }
  }
  // Keccak padding (0x01 after the message, 0x80 in the last byte of the
  // last block) and zero fill up to total_words.
  __device__ __forceinline__ void finish_keccak(uint32_t total_words) {
    push(0x01, 1);
    if (nw == total_words) {  // the 0x01 byte was the last byte of the block
      out[total_words - 1] |= 0x80ull << 56;
      return;
    }
    for (uint32_t w = nw; w < total_words; w++) {
      uint64_t v = w == nw ? acc : 0ull;
      if (w + 1 == total_words) v |= 0x80ull << 56;
      out[w] = v;
    }
  }
  // flush and zero fill up to total_words
  __device__ __forceinline__ void finish_raw(uint32_t total_words) {
    for (uint32_t w = nw; w < total_words; w++) out[w] = w == nw ? acc : 0ull;
  }
};

struct ChildRef {
  const uint8_t *p;
  uint32_t len;
};

__device__ __forceinline__ ChildRef child_bytes(const Pools &P, ref_t c) {
  ChildRef cr;
  uint32_t t = ref_type(c), i = ref_idx(c);
  if (t == T_VALUE) {
    const Value &V = P.value[i];
    cr.p = (const uint8_t *)V.addr;
    cr.len = V.size;
  } else if (t == T_BRANCH) {
    cr.p = P.branch[i].hash;
    cr.len = P.branch[i].hash_size;
  } else {
    cr.p = P.shrt[i].hash;
    cr.len = P.shrt[i].hash_size;
  }
  return cr;
}

__device__ __forceinline__ uint32_t child_item_len(const ChildRef &c) {
  return rlp_item_len(c.len, c.len == 1 ? c.p[0] : 0x80u);
}

// Emits the node's full encoding (list header, then items) into a writer.
template <typename W>
__device__ __forceinline__ void emit_node(const Pools &P, ref_t r, uint32_t payload, W &w) {
  uint32_t hn;
  uint64_t h = rlp_hdr(payload, 0xC0, 0xF7, &hn);
  w.inline_bytes(h, hn);
  if (ref_type(r) == T_SHORT) {
    const Short &S = P.shrt[ref_idx(r)];
    w.rlp_compact(P.keys + S.key_off, S.key_start, S.key_len - S.term, S.term);
    ChildRef c = child_bytes(P, S.child);
    w.rlp_mem(c.p, c.len);
  } else {
    const Branch &B = P.branch[ref_idx(r)];
#pragma unroll
    for (int s = 0; s < 17; s++) {
      ref_t c = B.child[s];
      if (c == REF_NULL) {
        w.inline_bytes(0x80, 1);
      } else {
        ChildRef cb = child_bytes(P, c);
        w.rlp_mem(cb.p, cb.len);
      }
    }
  }
}

__device__ __forceinline__ uint32_t node_payload(const Pools &P, ref_t r) {
  if (ref_type(r) == T_SHORT) {
    const Short &S = P.shrt[ref_idx(r)];
    uint32_t kc = (S.key_len - S.term) / 2 + 1;
    ChildRef c = child_bytes(P, S.child);
    return rlp_item_len(kc, 0) + child_item_len(c);
  }
  const Branch &B = P.branch[ref_idx(r)];
  uint32_t pay = 0;
#pragma unroll
  for (int s = 0; s < 17; s++) {
    ref_t c = B.child[s];
    pay += c == REF_NULL ? 1u : child_item_len(child_bytes(P, c));
  }
  return pay;
}

__device__ __forceinline__ uint8_t *node_hash(const Pools &P, ref_t r) {
  return ref_type(r) == T_SHORT ? P.shrt[ref_idx(r)].hash : P.branch[ref_idx(r)].hash;
}
__device__ __forceinline__ uint32_t node_hash_size(const Pools &P, ref_t r) {
  return ref_type(r) == T_SHORT ? P.shrt[ref_idx(r)].hash_size : P.branch[ref_idx(r)].hash_size;
}
__device__ __forceinline__ void node_done(const Pools &P, ref_t r, uint32_t hs) {
  if (ref_type(r) == T_SHORT) {
    P.shrt[ref_idx(r)].hash_size = (uint8_t)hs;
    P.shrt[ref_idx(r)].dirty = 0;
  } else {
    P.branch[ref_idx(r)].hash_size = (uint8_t)hs;
    P.branch[ref_idx(r)].dirty = 0;
  }
}

// Generic encoder of node r (loads interleaved with stores): enc_len[i], and
// the padded message in out (>= 32 bytes) or the raw encoding as the node's
// "hash" (< 32 bytes, cpu_mpt.cuh:413-418).  Returns the permutations needed.
__device__ __forceinline__ uint32_t encode_generic(const Pools &P, ref_t r, uint64_t *out, uint32_t *enc_out) {
  uint32_t payload = node_payload(P, r);
  uint32_t enc = rlp_hdr_len(payload) + payload;
  *enc_out = enc;
  if (enc < 32) {
    WordWriter w((uint64_t *)node_hash(P, r));
    emit_node(P, r, payload, w);
    w.finish_raw(4);
    node_done(P, r, enc);
    return 0;
  }
  uint32_t nb = keccak_blocks(enc);
  WordWriter w(out);
  emit_node(P, r, payload, w);
  w.finish_keccak(nb * 17);
  return nb;
}

// bytes [off, off + 8) of the little-endian 16-byte pair (a, b), off in 0..7
__device__ __forceinline__ uint64_t funnel8(uint64_t a, uint64_t b, uint32_t off) {
  return off ? (a >> (8 * off)) | (b << (64 - 8 * off)) : a;
}
__device__ __forceinline__ uint64_t low_bytes(uint64_t v, uint32_t n) { return n >= 8 ? v : v & ((1ull << (8 * n)) - 1); }


// Work counters of a hash pass: one reduction over the encoding lengths
// after the last level (per-wave atomics on one counter serialise when waves
// are many).
// (one partial per block: atomics from every block on four counters
// serialise)
__global__ __launch_bounds__(256) void k_hc_count(unsigned long long *parts, const uint32_t *__restrict__ enc_len,
                                                  const uint32_t *__restrict__ nvalid, uint32_t n_leaf) {
  __shared__ unsigned long long sh[4][4];
  const uint32_t n = *nvalid;
  // perms, bytes over all entries; perms, bytes over the leaf entries
  unsigned long long v[4] = {0, 0, 0, 0};
  // four entries per load (enc_len is 16-byte aligned)
  const uint32_t stride = 4 * gridDim.x * blockDim.x;
  for (uint32_t i0 = 4 * grid_tid(); i0 < n; i0 += stride) {
    uint32_t e4[4];
    if (i0 + 4 <= n) {
      const uint4 q = *(const uint4 *)(enc_len + i0);
      e4[0] = q.x, e4[1] = q.y, e4[2] = q.z, e4[3] = q.w;
    } else {
#pragma unroll
      for (uint32_t k = 0; k < 4; k++) e4[k] = i0 + k < n ? enc_len[i0 + k] : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
      const uint32_t e = e4[k];
      const uint32_t p = e >= 32 ? keccak_blocks(e) : 0u;
      v[0] += p;
      v[1] += e;
      if (i0 + k < n_leaf) {
        v[2] += p;
        v[3] += e;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; q++)
    for (int off = 32; off > 0; off >>= 1) v[q] += (unsigned long long)__shfl_xor((long long)v[q], off);
  if (__lane_id() == 0)
#pragma unroll
    for (int q = 0; q < 4; q++) sh[threadIdx.x >> 6][q] = v[q];
  __syncthreads();
  if (threadIdx.x < 4)
    parts[4 * blockIdx.x + threadIdx.x] =
        sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
}

// Writes the list header, the compact-key item and the value header (or a
// single raw value byte) of a leaf [compact key, value] from registers into
// `out` as whole little-endian words; the bytes past the head are zero.
__device__ __forceinline__ void leaf_head(uint64_t *out, uint32_t payload, const uint64_t kw[6], uint32_t koff,
                                          uint32_t ks, uint32_t cnt, uint32_t term, uint32_t vsz, bool v_raw,
                                          uint32_t vb0) {
  WordWriter w(out);
  uint32_t hn;
  uint64_t h = rlp_hdr(payload, 0xC0, 0xF7, &hn);
  w.inline_bytes(h, hn);
  const uint32_t kc = cnt / 2 + 1;
  uint64_t b0 = funnel8(kw[0], kw[1], koff);
  uint32_t first_nib = (ks & 1) ? (uint32_t)(b0 & 15) : (uint32_t)((b0 >> 4) & 15);
  uint32_t flag = (term << 5) | ((cnt & 1) << 4) | ((cnt & 1) ? first_nib : 0u);
  if (kc == 1) {
    w.inline_bytes(flag, 1);
  } else {
    w.inline_bytes(0x80 + kc, 1);
    w.inline_bytes(flag, 1);
    const uint32_t first = ks + (cnt & 1);
    const uint32_t d = first / 2 - ks / 2;  // 0 or 1
    uint64_t bs[5];
#pragma unroll
    for (int k = 0; k < 5; k++) bs[k] = funnel8(kw[k], kw[k + 1], koff);
    uint32_t c = kc - 1;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if ((uint32_t)(8 * k) < c) {
        uint64_t v;
        if (first & 1) {  // pairs straddle bytes
          uint64_t x = bs[k], y = (bs[k] >> 8) | (bs[k + 1] << 56);
          v = ((x << 4) & 0xF0F0F0F0F0F0F0F0ull) | ((y >> 4) & 0x0F0F0F0F0F0F0F0Full);
        } else {
          v = d ? (bs[k] >> 8) | (bs[k + 1] << 56) : bs[k];
        }
        uint32_t t = min(8u, c - 8 * k);
        w.push(low_bytes(v, t), t);
      }
    }
  }
  if (v_raw) {
    w.inline_bytes(vb0, 1);
  } else {
    h = rlp_hdr(vsz, 0x80, 0xB7, &hn);
    w.inline_bytes(h, hn);
  }
  if (w.accn) out[w.nw] = w.acc;
}

// Bytes [o, o + 8) of the value at va (vsz bytes), zero outside [0, vsz),
// little-endian; reads only aligned words that hold value bytes.
__device__ __forceinline__ uint64_t value_word(uint64_t va, uint32_t vsz, int64_t o) {
  const int64_t lo = o > 0 ? o : 0, hi = o + 8 < (int64_t)vsz ? o + 8 : (int64_t)vsz;
  if (lo >= hi) return 0ull;
  const uint64_t a_lo = va + (uint64_t)lo;
  const uint64_t *q = (const uint64_t *)(a_lo & ~7ull);
  const uint32_t sh = (uint32_t)(a_lo & 7);
  const uint32_t nbytes = (uint32_t)(hi - lo);
  uint64_t v = q[0] >> (8 * sh);
  if (sh + nbytes > 8) v |= q[1] << (64 - 8 * sh);
  v = low_bytes(v, nbytes);
  return v << (8 * (uint32_t)(lo - o));
}

constexpr uint32_t kLeafGroup = 8;                 // lanes per leaf
constexpr uint32_t kLeafNodes = 256 / kLeafGroup;  // leaves per block
constexpr uint32_t kLeafHeadWords = 6;             // list header + compact key + value header <= 44 bytes
constexpr uint32_t kLeafChunk = 5;                 // message words per lane per load/store round

// Leaf level: 8 lanes per leaf short node [compact key, value] whose key
// remainder has <= 64 nibbles.  Lane s computes message words s, s + 8, ...
// of the padded message -- head bytes, value bytes (two aligned loads per
// word, consecutive across the group) and padding -- and stores them, so the
// group's loads and stores are contiguous and no store waits on a later load.
// Lane 0 builds the head (list header, compact key, value header) in LDS.
// The dependent loads are kept short: the key words are loaded with the value
// node, the value's first byte only for a 1-byte value (which may be raw), and
// the first chunk of value words before the block waits for the heads.
// Other leaves take the generic encoder on lane 0.
__global__ __launch_bounds__(256) void k_hc_encode_leaf(Pools P, const ref_t *__restrict__ list,
                                                        const uint64_t *__restrict__ slot, uint32_t n,
                                                        uint64_t *__restrict__ scratch,
                                                        uint32_t *__restrict__ enc_len) {
  if (P.cnt->graph_bad) return;  // a replayed pass whose batch is unlike the captured one
  __shared__ uint64_t head[kLeafNodes][kLeafHeadWords];
  const uint32_t g = threadIdx.x / kLeafGroup, s = threadIdx.x % kLeafGroup;
  const uint32_t i = blockIdx.x * kLeafNodes + g;
  const bool active = i < n;
  const ref_t r = active ? list[i] : REF_NULL;
  uint64_t key_off = 0, va = 0;
  uint32_t ks = 0, cnt = 0, term = 0, vsz = 0;
  ref_t ch = REF_NULL;
  if (active) {
    const Short &S = P.shrt[ref_idx(r)];
    ch = S.child;
    key_off = S.key_off;
    ks = S.key_start;
    cnt = (uint32_t)S.key_len - S.term;
    term = S.term;
  }
  bool fast = ref_type(ch) == T_VALUE && cnt <= kLeafFastNibbles;
  // packed key bytes from byte ks / 2: up to 33 bytes, 6 aligned words
  uint64_t kw[6] = {0, 0, 0, 0, 0, 0};
  uint32_t koff = 0;
  if (fast && s == 0) {
    const uintptr_t ka = (uintptr_t)(P.keys + key_off + ks / 2);
    koff = (uint32_t)(ka & 7);
    const uint64_t *kw_p = (const uint64_t *)(ka - koff);
    const uint32_t kbytes = cnt ? (ks + cnt - 1) / 2 - ks / 2 + 1 : 1;
    const uint32_t nkw = (koff + kbytes + 1 + 7) / 8;
#pragma unroll
    for (int q = 0; q < 6; q++) kw[q] = (uint32_t)q < nkw ? kw_p[q] : 0ull;
  }
  if (fast) {
    const Value &V = P.value[ref_idx(ch)];
    va = V.addr;
    vsz = V.size;
  }
  const uint32_t vb0 = fast && vsz == 1 ? (uint32_t)(value_word(va, 1, 0) & 0xff) : 0u;
  const uint32_t kc = cnt / 2 + 1;
  const uint32_t key_item = kc == 1 ? 1u : 1u + kc;
  const bool v_raw = vsz == 1 && vb0 <= 0x7f;
  const uint32_t vh = v_raw ? 0u : rlp_hdr_len(vsz);
  const uint32_t payload = key_item + (v_raw ? 1u : vh + vsz);
  const uint32_t enc = rlp_hdr_len(payload) + payload;
  fast = fast && enc >= 32;
  const uint32_t nw = fast ? 17 * keccak_blocks(enc) : 0u;
  const uint32_t vstart = enc - vsz;  // message offset of the value bytes (unless raw, in the head)
  // this lane's words in chunks of kLeafChunk: every load of a chunk is
  // issued before its stores (a load after a store waits for the store)
  uint64_t word[kLeafChunk];
  auto chunk = [&](uint32_t w0) {
#pragma unroll
    for (uint32_t q = 0; q < kLeafChunk; q++) {
      const uint32_t w = w0 + q * kLeafGroup, x0 = 8 * w;
      uint64_t v = 0;
      if (w < nw && !v_raw && x0 + 8 > vstart && x0 < enc) v = value_word(va, vsz, (int64_t)x0 - (int64_t)vstart);
      if (enc >= x0 && enc < x0 + 8) v |= 0x01ull << (8 * (enc - x0));  // Keccak padding
      if (w + 1 == nw) v |= 0x80ull << 56;
      word[q] = v;
    }
  };
  chunk(s);
  if (fast && s == 0) {
#pragma unroll
    for (uint32_t q = 0; q < kLeafHeadWords; q++) head[g][q] = 0;
    leaf_head(head[g], payload, kw, koff, ks, cnt, term, vsz, v_raw, vb0);
  }
  __syncthreads();
  if (fast) {
    static_assert(kLeafHeadWords <= kLeafGroup, "head words are in every lane's first word");
    uint64_t *out = scratch + slot[i] / 8;
    if (s < kLeafHeadWords) word[0] |= head[g][s];
    for (uint32_t w0 = s;;) {
#pragma unroll
      for (uint32_t q = 0; q < kLeafChunk; q++)
        if (w0 + q * kLeafGroup < nw) out[w0 + q * kLeafGroup] = word[q];
      w0 += kLeafGroup * kLeafChunk;
      if (w0 >= nw) break;
      chunk(w0);
    }
    if (s == 0) enc_len[i] = enc;
  } else if (active && s == 0) {
    uint32_t e;
    encode_generic(P, r, scratch + slot[i] / 8, &e);
    enc_len[i] = e;
  }
}

// An extension (is_direct_ext) encodes like a leaf whose value is its child's
// hash buffer: [compact key, string of the child's digest, or of its raw
// encoding when that is under 32 bytes].  Fills that leaf's metadata words
// (the string's bytes are read from the child node) and returns true for
// such a node.
__device__ __forceinline__ bool ext_direct_meta(const Pools &P, ref_t r, uint64_t m[2]) {
  if (ref_type(r) != T_SHORT) return false;
  const Short &S = P.shrt[ref_idx(r)];
  if (!is_direct_ext(S)) return false;
  const Branch &B = P.branch[ref_idx(S.child)];
  leaf_meta_words(S.key_off, S.key_start, S.key_len, 0, B.hash_size, (uint64_t)(uintptr_t)B.hash, m);
  return true;
}

constexpr uint32_t kInnerGroup = 16;              // lanes per node (one per nibble slot)
constexpr uint32_t kInnerNodes = 256 / kInnerGroup;
constexpr uint32_t kInnerWords = 80;              // 640 LDS bytes per node >= 4 Keccak blocks

// Encodes inner node r with 16 lanes (s = 0..15 of a 16-lane-aligned segment,
// part = this thread is one of them).  For a branch whose 16 nibble slots hold
// only empty slots and hashed (32-byte) children, and whose slot 16 is empty
// (every branch of fixed-length keys), lane s loads child s's digest, the
// segment prefix-sums the item lengths, each lane writes its item into the
// node's LDS copy of the padded message (lds, kInnerWords words), and the
// segment streams the message words to out.  Other nodes (extension short
// nodes, branches with embedded children or a value) take the generic encoder
// on lane 0.  Every thread of the block calls this (it synchronises the
// block); lane 0 returns the encoding length.  With in_lds, the LDS copy of
// a fast node's message stands (*in_lds set on the segment's lanes) and out
// is not written; the generic encoder still writes out.
__device__ __forceinline__ uint32_t encode_inner_node(const Pools &P, ref_t r, bool active, bool part, uint32_t s,
                                                      uint64_t *lds, uint64_t *out, bool *in_lds = nullptr) {
  const bool is_branch = part && active && ref_type(r) == T_BRANCH;
  ref_t c = REF_NULL, c16 = REF_NULL;
  if (is_branch) {
    const Branch &B = P.branch[ref_idx(r)];
    c = B.child[s];
    c16 = B.child[16];
  }
  const uint32_t ct = ref_type(c);
  const bool node_child = ct == T_BRANCH || ct == T_SHORT;
  uint64_t h[4] = {0, 0, 0, 0};
  uint32_t hs = 0;
  if (node_child) {
    const uint64_t *hp = (const uint64_t *)node_hash(P, c);
#pragma unroll
    for (int q = 0; q < 4; q++) h[q] = hp[q];
    hs = node_hash_size(P, c);
  }
  const bool ok = c == REF_NULL || (node_child && hs == 32);
  const uint64_t bad = __ballot(!ok);
  const uint32_t gbase = __lane_id() & ~(kInnerGroup - 1);
  const bool fast = is_branch && c16 == REF_NULL && ((bad >> gbase) & 0xffffull) == 0;
  if (part) {
#pragma unroll
    for (uint32_t q = 0; q < kInnerWords / kInnerGroup; q++) lds[s + kInnerGroup * q] = 0;
  }
  __syncthreads();
  uint32_t enc = 0, nb = 0;
  if (fast) {
    const uint32_t len = c == REF_NULL ? 1u : 33u;
    uint32_t incl = len;
#pragma unroll
    for (uint32_t d = 1; d < kInnerGroup; d <<= 1) {
      uint32_t t = (uint32_t)__shfl_up((int)incl, d, kInnerGroup);
      if (s >= d) incl += t;
    }
    const uint32_t total = (uint32_t)__shfl((int)incl, kInnerGroup - 1, kInnerGroup);
    const uint32_t payload = total + 1;  // slot 16 is empty: one 0x80 byte
    uint32_t hn;
    const uint64_t hdr = rlp_hdr(payload, 0xC0, 0xF7, &hn);
    enc = hn + payload;
    nb = keccak_blocks(enc);
    uint8_t *b = (uint8_t *)lds;
    const uint32_t pos = hn + incl - len;
    if (c == REF_NULL) {
      b[pos] = 0x80;
    } else {
      b[pos] = 0xa0;
#pragma unroll
      for (int q = 0; q < 32; q++) b[pos + 1 + q] = (uint8_t)(h[q >> 3] >> (8 * (q & 7)));
    }
    if (s == 0) {
      for (uint32_t q = 0; q < hn; q++) b[q] = (uint8_t)(hdr >> (8 * q));
      b[enc - 1] = 0x80;  // slot 16
      b[enc] = 0x01;      // Keccak padding
      b[136 * nb - 1] |= 0x80;
    }
  }
  __syncthreads();
  if (in_lds) *in_lds = fast;
  if (fast) {
    if (!in_lds)
      for (uint32_t w = s; w < 17 * nb; w += kInnerGroup) out[w] = lds[w];
  } else if (part && active && s == 0) {
    encode_generic(P, r, out, &enc);
  }
  return enc;
}

// Inner levels: 16 lanes per node.
__global__ __launch_bounds__(256) void k_hc_encode_inner(Pools P, const ref_t *__restrict__ list,
                                                         const uint64_t *__restrict__ slot, uint32_t n,
                                                         uint64_t *__restrict__ scratch,
                                                         uint32_t *__restrict__ enc_len) {
  if (P.cnt->graph_bad) return;  // a replayed pass whose batch is unlike the captured one
  __shared__ uint64_t lds[kInnerNodes][kInnerWords];
  const uint32_t g = threadIdx.x / kInnerGroup, s = threadIdx.x % kInnerGroup;
  const uint32_t i = blockIdx.x * kInnerNodes + g;
  const bool active = i < n;
  const ref_t r = active ? list[i] : REF_NULL;
  const uint32_t enc = encode_inner_node(P, r, active, true, s, lds[g], active ? scratch + slot[i] / 8 : scratch);
  if (active && s == 0) enc_len[i] = enc;
}

// ----------------------------------------------------------------- permute --

// Sponge over nb padded blocks with one state per lane.
__device__ __forceinline__ void keccak_lane(const uint64_t *m, uint32_t nb, uint64_t s[25]) {
#pragma unroll
  for (int q = 0; q < 25; q++) s[q] = 0;
  for (uint32_t b = 0; b < nb; b++) {
#pragma unroll
    for (int q = 0; q < 17; q++) s[q] ^= m[17 * b + q];
    keccak_f1600(s);
  }
}

// Writes the item 0xa0 + digest (33 bytes; digest words d0, d1 as stored)
// into a zeroed LDS row at byte offset off.  With r = off mod 4 the item
// covers bytes [r, r + 33) of dwords off / 4 + 0 .. 8: dwords 1 .. 7 lie
// inside it and are stored plainly; dwords 0 and 8 may hold a neighbour's
// bytes (the previous or next item, an empty slot's 0x80, the list header)
// and are ORed in, so neighbouring children need no ordering.
__device__ __forceinline__ void or_child_item(uint32_t *row32, uint32_t off, const uint4 &d0, const uint4 &d1) {
  // the item's dwords (byte 0 = 0xa0, bytes 1 .. 32 = the digest)
  uint32_t it[9];
  it[0] = 0xa0u | (d0.x << 8);
  it[1] = __builtin_amdgcn_alignbyte(d0.y, d0.x, 3);
  it[2] = __builtin_amdgcn_alignbyte(d0.z, d0.y, 3);
  it[3] = __builtin_amdgcn_alignbyte(d0.w, d0.z, 3);
  it[4] = __builtin_amdgcn_alignbyte(d1.x, d0.w, 3);
  it[5] = __builtin_amdgcn_alignbyte(d1.y, d1.x, 3);
  it[6] = __builtin_amdgcn_alignbyte(d1.z, d1.y, 3);
  it[7] = __builtin_amdgcn_alignbyte(d1.w, d1.z, 3);
  it[8] = d1.w >> 24;
  const uint32_t r = off & 3u, w0 = off >> 2;
  uint32_t *p = row32 + w0;
  // shifted by r bytes: dword m takes item bytes [4m - r, 4m - r + 4)
#pragma unroll
  for (int m = 0; m < 9; m++) {
    const uint32_t prev = m ? it[m - 1] : 0u;
    const uint32_t v = r ? __builtin_amdgcn_alignbyte(it[m], prev, 4 - r) : it[m];
    if (m == 0 || m == 8)
      atomicOr(p + m, v);
    else
      p[m] = v;
  }
}

// Inner levels above the spread threshold whose nodes are branches of at
// most NB Keccak blocks and extensions (all but a sixteenth of them),
// encoded and hashed in one launch with the messages in LDS: one wave per
// block, one node per lane for the permutation, but the encoding's memory
// work spread over the whole wave:
//  1. Each lane reads its branch's 16 child references (one 64-byte load),
//     keeps them in LDS and writes the bytes that depend on the branch alone
//     into its LDS row: list header, 0x80 for each empty slot, 0x80 for slot
//     16, Keccak padding.  An extension (a short node over a branch) is
//     written whole by its lane: [compact key, 0xa0 + the child's digest].
//  2. The wave's children are dealt out to the lanes: for rows of 1 or 2
//     blocks (a few children per node) the present children, numbered by a
//     prefix sum of the lanes' child counts, round robin; for rows of 4
//     blocks all 64 x 16 (node, slot) pairs, 16 per lane, four at a time so
//     their loads overlap.  Each present child's digest is loaded and ORed,
//     as 0xa0 + digest, into its parent's row at the child's offset (the
//     rank of its slot among the node's present children).
//  3. Each lane absorbs its row and permutes.
// A node whose message would be taller than the rows (NB blocks), a branch
// with a value in slot 16 or an embedded child, and any other short node
// take the generic encoder and their scratch slot.  Blocks run in reverse
// order: the level's nodes are sorted by block count with the extensions
// last, so the larger nodes and extensions start first and finish under the
// bulk.  Entries [skip_lo, skip_hi), the level's nodes taller than the rows
// (classes above NB blocks, before the extensions), are hashed two per block
// by the launch's first nx blocks in the spread layout (hash_pair_spread2):
// dispatched first, they run beside the bulk instead of holding the launch
// open after it.
__device__ __forceinline__ void hash_pair_spread2(const Pools &P, const ref_t *__restrict__ list,
                                                  const uint64_t *__restrict__ slot, uint32_t n,
                                                  uint64_t *__restrict__ scratch, uint32_t *__restrict__ enc_len,
                                                  uint32_t pair, uint64_t (*lds)[kInnerWords]);
template <int NB>
__global__ __launch_bounds__(64) void k_hc_inner_wave(Pools P, const ref_t *__restrict__ list,
                                                      const uint64_t *__restrict__ slot, uint32_t n,
                                                      uint64_t *__restrict__ scratch, uint32_t *__restrict__ enc_len,
                                                      uint32_t skip_lo, uint32_t skip_hi, uint32_t nx) {
  if (P.cnt->graph_bad) return;  // a replayed pass whose batch is unlike the captured one
  constexpr uint32_t RS = 17 * NB + 1;  // row stride in words (odd: lane-strided reads spread over the banks)
  __shared__ uint64_t rows[64 * RS];
  static_assert(64 * RS >= 2 * kInnerWords, "a pair's messages fit the rows");
  if (blockIdx.x < nx) {  // uniform over the block
    __builtin_amdgcn_s_setprio(2);  // latency-bound waves beside the bulk's permutations
    hash_pair_spread2(P, list + skip_lo, slot + skip_lo, skip_hi - skip_lo, scratch, enc_len + skip_lo, blockIdx.x,
                      (uint64_t(*)[kInnerWords])rows);
    return;
  }
  constexpr bool kDense = NB >= 4;      // many children per node: (node, slot) pairs dealt out directly
  __shared__ ref_t cref[kDense ? 64 * 16 : 1];  // the lanes' branches' child references (dense dealing)
  __shared__ ref_t nref[64];            // the lanes' nodes
  __shared__ uint32_t nbase[64];        // first child number of each node (compact dealing)
  __shared__ uint32_t ninfo[64];        // child mask (0 unless the row path encodes the node) | header length << 16
  __shared__ uint32_t bad[64];          // 1: the node takes the generic encoder
  const uint32_t lane = threadIdx.x;
  const uint32_t i = (gridDim.x - 1 - blockIdx.x) * 64 + lane;
  const bool active = i < n && (i < skip_lo || i >= skip_hi);
  const ref_t r = active ? list[i] : REF_NULL;
  const bool branch = active && ref_type(r) == T_BRANCH;
  bool fast = branch, ext = false;
  uint32_t mask = 0, enc = 0, nb = 0, hn = 0;
  uint64_t hdr = 0;
  if (branch) {
    const Branch &B = P.branch[ref_idx(r)];
    const uint4 *cw = (const uint4 *)B.child;  // child[0 .. 15]: 64 bytes, 16-byte aligned
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint4 v = cw[q];
      if (kDense) ((uint4 *)(cref + 16 * lane))[q] = v;
      mask |= (v.x != REF_NULL ? 1u : 0u) << (4 * q);
      mask |= (v.y != REF_NULL ? 2u : 0u) << (4 * q);
      mask |= (v.z != REF_NULL ? 4u : 0u) << (4 * q);
      mask |= (v.w != REF_NULL ? 8u : 0u) << (4 * q);
    }
    const uint32_t k = __popc(mask);
    const uint32_t payload = 33 * k + (16 - k) + 1;
    hdr = rlp_hdr(payload, 0xC0, 0xF7, &hn);
    enc = hn + payload;
    nb = keccak_blocks(enc);
    fast = B.child[16] == REF_NULL && nb <= (uint32_t)NB && enc >= 32;
  } else if (active && ref_type(r) == T_SHORT) {
    ext = is_direct_ext(P.shrt[ref_idx(r)]);
  }
  const uint32_t kc = fast ? (uint32_t)__popc(mask) : 0u;
  const uint32_t incl = wave_inclusive_scan(kc);
  const uint32_t total = (uint32_t)__shfl((int)incl, 63, 64);  // the wave's row-path children
  nref[lane] = r;
  nbase[lane] = incl - kc;
  ninfo[lane] = (fast ? mask : 0u) | (hn << 16);
  bad[lane] = fast ? 0u : 1u;
  uint64_t *row = rows + lane * RS;
  if (fast)
    for (uint32_t w = 0; w < 17 * nb; w++) row[w] = 0;
  __syncthreads();  // rows zeroed; child references, masks visible to the wave
  if (fast) {
    uint8_t *rb = (uint8_t *)row;
    for (uint32_t q = 0; q < hn; q++) rb[q] = (uint8_t)(hdr >> (8 * q));
    uint32_t rank = 0;  // present children below slot s
#pragma unroll
    for (uint32_t s = 0; s < 16; s++) {
      const bool has = (mask >> s) & 1u;
      if (!has) rb[hn + 33 * rank + (s - rank)] = 0x80;
      rank += has ? 1u : 0u;
    }
    rb[enc - 1] = 0x80;  // slot 16: empty
    rb[enc] = 0x01;      // Keccak padding; enc < 136 * nb, so this byte is the message's own
    rb[136 * nb - 1] |= 0x80;
  } else if (ext) {
    // [compact key, string of the child's digest]: the whole row from this lane
    const Short &S = P.shrt[ref_idx(r)];
    const Branch &C = P.branch[ref_idx(S.child)];
    const uint4 d0 = ((const uint4 *)C.hash)[0], d1 = ((const uint4 *)C.hash)[1];
    const uint32_t cnt = (uint32_t)S.key_len - S.term;
    const uint32_t payload = rlp_item_len(cnt / 2 + 1, 0) + 33;
    uint32_t ehn;
    const uint64_t ehdr = rlp_hdr(payload, 0xC0, 0xF7, &ehn);
    const uint32_t eenc = ehn + payload;
    if (C.hash_size == 32 && keccak_blocks(eenc) <= (uint32_t)NB) {
      WordWriter w(row);
      w.inline_bytes(ehdr, ehn);
      w.rlp_compact(P.keys + S.key_off, S.key_start, cnt, S.term);
      w.inline_bytes(0xa0, 1);
      w.push(((uint64_t)d0.y << 32) | d0.x, 8);
      w.push(((uint64_t)d0.w << 32) | d0.z, 8);
      w.push(((uint64_t)d1.y << 32) | d1.x, 8);
      w.push(((uint64_t)d1.w << 32) | d1.z, 8);
      nb = keccak_blocks(eenc);
      w.finish_keccak(17 * nb);
      enc = eenc;
      bad[lane] = 0;
      fast = true;
    }
  }
  if (!kDense) {
    // the wave's present children, numbered node by node, dealt out round
    // robin (a few per lane)
    for (uint32_t t = lane; t < total; t += 64) {
      uint32_t lo = 0, hi = 63;  // the node: the last with nbase <= t (it has children)
      while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (nbase[mid] <= t)
          lo = mid;
        else
          hi = mid - 1;
      }
      const uint32_t j = lo, info = ninfo[j];
      const uint32_t o = t - nbase[j];  // the child's rank among the node's present children
      uint32_t mm = info & 0xffffu;     // its slot: the o-th set bit of the mask
      for (uint32_t q = 0; q < o; q++) mm &= mm - 1;
      const uint32_t sl = __ffs(mm) - 1;
      const ref_t c = P.branch[ref_idx(nref[j])].child[sl];
      const bool sh = ref_type(c) == T_SHORT;
      const uint8_t *h = sh ? P.shrt[ref_idx(c)].hash : P.branch[ref_idx(c)].hash;
      const uint4 d0 = ((const uint4 *)h)[0], d1 = ((const uint4 *)h)[1];
      const uint32_t hs = sh ? P.shrt[ref_idx(c)].hash_size : P.branch[ref_idx(c)].hash_size;
      if (hs != 32) {
        bad[j] = 1;  // an embedded child: the node's own lane encodes it generically
        continue;
      }
      or_child_item((uint32_t *)(rows + j * RS), (info >> 16) + 33 * o + (sl - o), d0, d1);
    }
  }
  // (dense) the wave's (node, slot) pairs: pair p = node p / 16, slot p % 16;
  // lane l takes pairs l + 64 q, four at a time so their digest loads overlap
#pragma unroll 1
  for (uint32_t q0 = 0; kDense && q0 < 16; q0 += 4) {
    uint4 d0[4], d1[4];
    uint32_t hs[4], off[4], jj[4];
    bool has[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const uint32_t p = lane + 64 * (q0 + u);
      const uint32_t j = p >> 4, s = p & 15u;
      const uint32_t info = ninfo[j];
      const uint32_t m = info & 0xffffu;
      const uint32_t o = __popc(m & ((1u << s) - 1u));  // the child's rank among the node's present children
      has[u] = (m >> s) & 1u;
      jj[u] = j;
      off[u] = (info >> 16) + 33 * o + (s - o);
      hs[u] = 0;
      d0[u] = d1[u] = make_uint4(0, 0, 0, 0);
      if (has[u]) {
        const ref_t c = cref[16 * j + s];
        const bool sh = ref_type(c) == T_SHORT;
        const uint8_t *h = sh ? P.shrt[ref_idx(c)].hash : P.branch[ref_idx(c)].hash;
        d0[u] = ((const uint4 *)h)[0];
        d1[u] = ((const uint4 *)h)[1];
        hs[u] = sh ? P.shrt[ref_idx(c)].hash_size : P.branch[ref_idx(c)].hash_size;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (!has[u]) continue;
      if (hs[u] != 32) {
        bad[jj[u]] = 1;  // an embedded child: the node's own lane encodes it generically
        continue;
      }
      or_child_item((uint32_t *)(rows + jj[u] * RS), off[u], d0[u], d1[u]);
    }
  }
  __syncthreads();  // every row complete
  if (!active) return;
  const bool use_row = fast && bad[lane] == 0;
  uint64_t *out = scratch + slot[i] / 8;
  if (!use_row) {
    encode_generic(P, r, out, &enc);
    nb = enc >= 32 ? keccak_blocks(enc) : 0u;
  }
  enc_len[i] = enc;
  if (enc < 32) return;  // the raw encoding is the node's "hash", stored by the encoder
  uint64_t st[25];
#pragma unroll
  for (int q = 0; q < 25; q++) st[q] = 0;
  for (uint32_t b = 0; b < nb; b++) {
    if (use_row) {
#pragma unroll
      for (int q = 0; q < 17; q++) st[q] ^= row[17 * b + q];
    } else {
#pragma unroll
      for (int q = 0; q < 17; q++) st[q] ^= out[17 * b + q];
    }
    keccak_f1600(st);
  }
  uint64_t *hp = (uint64_t *)node_hash(P, r);
  hp[0] = st[0];
  hp[1] = st[1];
  hp[2] = st[2];
  hp[3] = st[3];
  node_done(P, r, 32);
}

__global__ __launch_bounds__(256) void k_hc_perm_lane(Pools P, const ref_t *list, const uint64_t *slot,
                                                      const uint32_t *enc_len, uint32_t n,
                                                      const uint64_t *scratch) {
  if (P.cnt->graph_bad) return;  // a replayed pass whose batch is unlike the captured one
  uint32_t i = grid_tid();
  if (i >= n) return;
  uint32_t enc = enc_len[i];
  if (enc < 32) return;
  uint64_t s[25];
  keccak_lane(scratch + slot[i] / 8, keccak_blocks(enc), s);
  ref_t r = list[i];
  uint64_t *h = (uint64_t *)node_hash(P, r);
  h[0] = s[0];
  h[1] = s[1];
  h[2] = s[2];
  h[3] = s[3];
  node_done(P, r, 32);
}

// Leaf level 0 (key remainder <= 64 nibbles, value < 16,383 bytes, any
// number of Keccak blocks): encode and hash in one launch, one thread per
// leaf, without a scratch message.  The leaf's
// metadata comes from the classification (leaf_meta, sorted order), so the
// loads are at most two deep; the thread builds the message head (list
// header, compact key, value header) once, then computes each 8-byte message
// word as it absorbs it: head word, value bytes (aligned value words shifted
// by the value's offset in the message and masked to the value's bytes),
// Keccak padding.  The state stays in registers as in k_hc_perm_lane.
// The kernel reads its leaf count (level 0's bins) from the classification
// on the device, so it can be launched before the host has read the bins
// back.  Other leaves are binned at level 1; the only leaves here off the
// word builder's path are those whose encoding is under 32 bytes (raw
// "hash", no permutation).

// Aligned 8-byte load from global memory.
typedef const __attribute__((address_space(1))) uint64_t gword_t;
__device__ __forceinline__ uint64_t gload(uint64_t a) { return *(gword_t *)a; }

// Message layout of a leaf from its metadata word m0 (leaf_meta) and the
// value's first byte (used only for a 1-byte value, which may be raw).
struct LeafLayout {
  uint32_t cnt, vsz, payload, enc, nb;
  bool raw, fast;
};
__device__ __forceinline__ LeafLayout leaf_layout(uint64_t m0, uint32_t vb0) {
  LeafLayout l;
  l.cnt = (uint32_t)(m0 >> 41) & 255u;
  l.vsz = (uint32_t)(m0 >> 50);
  const uint32_t kc = l.cnt / 2 + 1;
  const uint32_t key_item = kc == 1 ? 1u : 1u + kc;
  l.raw = l.vsz == 1 && vb0 <= 0x7f;
  l.payload = key_item + (l.raw ? 1u : rlp_hdr_len(l.vsz) + l.vsz);
  l.enc = rlp_hdr_len(l.payload) + l.payload;
  l.nb = keccak_blocks(l.enc);
  l.fast = l.cnt <= kLeafFastNibbles && l.enc >= 32 && l.vsz <= kLeafDirectMaxValue;
  return l;
}

__device__ __forceinline__ int32_t clamp_i(int32_t x, int32_t lo, int32_t hi) { return max(lo, min(x, hi)); }

// Bytes [sh, sh + 8) of the little-endian 16-byte pair (lo, hi), sh in 0..7.
__device__ __forceinline__ uint64_t funnel_bytes(uint64_t lo, uint64_t hi, uint32_t sh) {
  const uint32_t d0 = (uint32_t)lo, d1 = (uint32_t)(lo >> 32), d2 = (uint32_t)hi, d3 = (uint32_t)(hi >> 32);
  const bool s4 = sh >= 4;
  const uint32_t x = s4 ? d1 : d0, y = s4 ? d2 : d1, z = s4 ? d3 : d2;
  const uint32_t r0 = __builtin_amdgcn_alignbyte(y, x, sh & 3);
  const uint32_t r1 = __builtin_amdgcn_alignbyte(z, y, sh & 3);
  return ((uint64_t)r1 << 32) | r0;
}

// A leaf whose encoding is under 32 bytes: the generic encoder writes the raw
// encoding as its hash.  Returns the encoding length.  Kept out of line: it
// is rare and large.
__device__ __noinline__ uint32_t leaf_raw(const Pools &P, ref_t r) {
  const uint32_t pay = node_payload(P, r);
  const uint32_t e = rlp_hdr_len(pay) + pay;
  if (e >= 32) {  // the binning keeps such leaves out of this kernel
    raise_error(P.cnt, ERR_INTERNAL, 5);
    return e;
  }
  uint32_t enc;
  encode_generic(P, r, nullptr, &enc);
  return enc;
}

constexpr uint32_t kLeafDirectHead = 7;  // LDS words per thread for the head (6 used; odd spreads banks)

// Message words [wlo, whi] that can hold a leaf's string bytes or its 0x01
// padding byte, from its metadata word m0 alone (a superset: the encoding
// length is taken as if the string had a header, the string's start one byte
// earlier than that).
__device__ __forceinline__ void leaf_value_words(uint64_t m0, uint32_t &wlo, uint32_t &whi) {
  const uint32_t cnt = (uint32_t)(m0 >> 41) & 255u, vsz = (uint32_t)(m0 >> 50);
  const uint32_t kc = cnt / 2 + 1;
  const uint32_t payload = (kc == 1 ? 1u : 1u + kc) + rlp_hdr_len(vsz) + vsz;
  const uint32_t enc = rlp_hdr_len(payload) + payload;
  wlo = (enc - vsz - 1) / 8;
  whi = enc / 8;
}

// Encodes and hashes one node shaped [compact key, string] from its metadata
// words m0 (leaf_meta_words) and va (the string's bytes), building each
// message word in registers; hp is the thread's LDS row for the message head.
// Only message words [wlo, whi] (the same on every lane of the wave, and
// covering each lane's leaf_value_words) are built from the string: the
// others hold head bytes or zeros.  Returns the encoding length.
__device__ __forceinline__ uint32_t hash_leaf_shaped(const Pools &P, ref_t r, uint64_t m0, uint64_t va, uint64_t *hp,
                                                     uint32_t wlo = 0, uint32_t whi = 1u << 30) {
  const uint32_t vsz = (uint32_t)(m0 >> 50);
  const uint32_t vb0 = vsz == 1 ? (uint32_t)(gload(va & ~7ull) >> (8 * (va & 7))) & 0xffu : 0u;
  const LeafLayout l = leaf_layout(m0, vb0);
  if (!l.fast) return leaf_raw(P, r);
  const uint32_t enc = l.enc, nb = l.nb;
  uint64_t hw[kLeafHeadWords] = {0, 0, 0, 0, 0, 0};
  {
    const uint64_t ka = (uint64_t)(uintptr_t)P.keys + (m0 & ((1ull << 40) - 1));
    const uint32_t koff = (uint32_t)(ka & 7);
    const uint32_t ks1 = (uint32_t)(m0 >> 40) & 1u, term = (uint32_t)(m0 >> 49) & 1u;
    const uint32_t kbytes = l.cnt ? (ks1 + l.cnt - 1) / 2 + 1 : 1;
    const uint32_t nkw = (koff + kbytes + 1 + 7) / 8;
    uint64_t kw[6];
#pragma unroll
    for (int q = 0; q < 6; q++) kw[q] = (uint32_t)q < nkw ? gload(ka - koff + 8 * q) : 0ull;
#pragma unroll
    for (uint32_t q = 0; q < kLeafHeadWords; q++) hp[q] = 0;
    leaf_head(hp, l.payload, kw, koff, ks1, l.cnt, term, vsz, l.raw, vb0);
#pragma unroll
    for (uint32_t q = 0; q < kLeafHeadWords; q++) hw[q] = hp[q];
  }
  // value geometry: message byte x is value byte x - vstart, i.e. byte
  // x + t0 from the aligned base V0; word w takes aligned words
  // jbase + w and jbase + w + 1 shifted by sh bytes
  const bool hasv = !l.raw && vsz > 0;
  const uint32_t vstart = enc - vsz;
  const uint64_t V0 = va & ~7ull;
  const uint32_t u = (uint32_t)(va & 7);
  const int32_t J = hasv ? (int32_t)((u + vsz - 1) >> 3) : 0;  // last aligned word holding value bytes
  const int32_t t0 = (int32_t)u - (int32_t)vstart;
  const uint32_t sh = (uint32_t)t0 & 7u;
  const int32_t jbase = (t0 - (int32_t)sh) >> 3;
  // the head lies in the first block, on bytes no value byte or padding
  // shares: the state starts as it
  uint64_t st[25];
#pragma unroll
  for (int q = 0; q < 25; q++) st[q] = q < (int)kLeafHeadWords ? hw[q] : 0ull;
#pragma unroll 1
  for (uint32_t b = 0; b < nb; b++) {
    {
      // this block's words of the window (wave-uniform bounds: the word
      // loops below branch over the words outside it)
      const int32_t qlo = max((int32_t)wlo - (int32_t)(17 * b), 0);
      const int32_t qhi = min((int32_t)whi - (int32_t)(17 * b), 16);
      uint64_t W[18];
#pragma unroll
      for (int q = 0; q < 18; q++) {
        W[q] = 0;
        if (q >= qlo && q <= qhi + 1) {
          const int32_t j = clamp_i(jbase + (int32_t)(17 * b) + q, 0, J);  // clamped into the value
          W[q] = hasv ? gload(V0 + 8 * (uint64_t)(uint32_t)j) : 0ull;
        }
      }
#pragma unroll
      for (int q = 0; q < 17; q++) {
        if (q < qlo || q > qhi) continue;
        const int32_t w = 17 * (int32_t)b + q;
        // value bytes: message bytes [max(vstart, 8w), min(enc, 8w + 8)) of this word
        const int32_t lo_b = clamp_i((int32_t)vstart - 8 * w, 0, 8);
        const int32_t hi_b = clamp_i((int32_t)enc - 8 * w, 0, 8);
        const uint64_t f = funnel_bytes(W[q], W[q + 1], sh);
        const uint64_t mk = hasv && hi_b > lo_b ? (~0ull << (8 * lo_b)) & (~0ull >> (64 - 8 * hi_b)) : 0ull;
        uint64_t v = f & mk;
        if ((int32_t)(enc >> 3) == w) v |= 0x01ull << (8 * (enc & 7));  // Keccak padding
        st[q] ^= v;
      }
      if (b + 1 == nb) st[16] ^= 0x80ull << 56;  // the last byte of the last block
    }
    keccak_f1600(st);
  }
  uint64_t *h = (uint64_t *)node_hash(P, r);
  h[0] = st[0];
  h[1] = st[1];
  h[2] = st[2];
  h[3] = st[3];
  node_done(P, r, 32);
  return enc;
}

__global__ __launch_bounds__(256) void k_hc_leaf_direct(Pools P, const ref_t *__restrict__ list,
                                                        const uint64_t *__restrict__ lmeta,
                                                        const uint32_t *__restrict__ hist,
                                                        uint32_t *__restrict__ enc_len) {
  if (P.cnt->graph_bad) return;  // a replayed pass whose batch is unlike the captured one
  __shared__ uint64_t head[256][kLeafDirectHead];
  const uint32_t j = grid_tid();
  uint32_t n = 0;  // the leaves of level 0, sorted first
#pragma unroll
  for (uint32_t c = 0; c < kLdsLeafBins; c++) n += hist[c];
  // the leaves from the last: the level is sorted by block count, so the
  // longest leaves start first and the 1-block ones fill the end of the launch
  const bool act = j < n;
  const uint32_t i = act ? n - 1 - j : 0u;
  const uint64_t m0 = act ? lmeta[2 * (size_t)i] : 0ull;
  // the wave's window of string words, reduced while every lane is active
  uint32_t wlo = 0xffffu, whi = 0;
  if (act) leaf_value_words(m0, wlo, whi);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    wlo = min(wlo, (uint32_t)__shfl_xor((int)wlo, off));
    whi = max(whi, (uint32_t)__shfl_xor((int)whi, off));
  }
  wlo = __builtin_amdgcn_readfirstlane(wlo);
  whi = __builtin_amdgcn_readfirstlane(whi);
  if (!act) return;  // no block barrier below
  enc_len[i] = hash_leaf_shaped(P, list[i], m0, lmeta[2 * (size_t)i + 1], head[threadIdx.x], wlo, whi);
}

// The leaf children of list[0 .. n) (inner nodes of the deepest levels),
// one thread per (node, slot): hashed here, on a second stream, so that
// those levels can be hashed beside the leaf-level launch, which hashes the
// same leaves again and stores the same bytes (a reader sees equal old and
// new values).  A child the word builder cannot take is a clean leaf, hashed
// by an earlier pass (the host runs no tail while such a leaf is dirty), and
// is left as it is.
__global__ __launch_bounds__(256) void k_hc_tail_leaves(Pools P, const ref_t *__restrict__ list, uint32_t n) {
  if (P.cnt->graph_bad) return;  // a replayed pass whose batch is unlike the captured one
  __builtin_amdgcn_s_setprio(2);  // ahead of the leaf-level launch's waves on a shared SIMD
  __shared__ uint64_t head[256][kLeafDirectHead];
  __shared__ ref_t job[256];
  __shared__ uint32_t sh[4];
  const uint32_t t = grid_tid(), i = t / 16, s = t % 16;
  ref_t leaf = REF_NULL;
  if (i < n) {
    const ref_t r = list[i];
    if (ref_type(r) == T_BRANCH) {  // an extension's child is a branch
      const ref_t c = P.branch[ref_idx(r)].child[s];
      if (ref_type(c) == T_SHORT && ref_type(P.shrt[ref_idx(c)].child) == T_VALUE) leaf = c;
    }
  }
  // the block's leaves packed into its first lanes (a few per node): the
  // waves left without one exit
  uint32_t tot;
  const uint32_t pos = block_exclusive_scan_256<uint32_t>(leaf != REF_NULL ? 1u : 0u, sh, &tot);
  if (leaf != REF_NULL) job[pos] = leaf;
  __syncthreads();
  if (threadIdx.x >= tot) return;  // no block barrier below
  const ref_t c = job[threadIdx.x];
  uint64_t m[2];
  leaf_meta(P, c, m);
  if (leaf_too_long(m[0])) return;  // clean: its hash stands
  (void)hash_leaf_shaped(P, c, m[0], m[1], head[threadIdx.x]);
}

// The extension class of an inner level (kExtClass): each extension encodes
// like a leaf (ext_direct_meta), so the leaf word builder hashes it with no
// scratch message.  A node that is not such an extension (none is binned
// here) would take the generic encoder and its scratch slot.
__global__ __launch_bounds__(256) void k_hc_ext_direct(Pools P, const ref_t *__restrict__ list,
                                                       const uint64_t *__restrict__ slot, uint32_t n,
                                                       uint64_t *__restrict__ scratch,
                                                       uint32_t *__restrict__ enc_len) {
  if (P.cnt->graph_bad) return;  // a replayed pass whose batch is unlike the captured one
  __shared__ uint64_t head[256][kLeafDirectHead];
  const uint32_t i = grid_tid();
  if (i >= n) return;  // no block barrier below
  const ref_t r = list[i];
  uint64_t m[2];
  if (ext_direct_meta(P, r, m)) {
    enc_len[i] = hash_leaf_shaped(P, r, m[0], m[1], head[threadIdx.x]);
    return;
  }
  uint32_t enc;
  uint64_t *msg = scratch + slot[i] / 8;
  encode_generic(P, r, msg, &enc);
  enc_len[i] = enc;
  if (enc < 32) return;
  uint64_t st[25];
  keccak_lane(msg, keccak_blocks(enc), st);
  uint64_t *h = (uint64_t *)node_hash(P, r);
  h[0] = st[0];
  h[1] = st[1];
  h[2] = st[2];
  h[3] = st[3];
  node_done(P, r, 32);
}

// Keccak-f[1600] with one state spread over a wave: lane 8x + y holds word
// A[x, y] (x, y < 5) as (l, h) halves; lanes 8x + 5 .. 8x + 7 and 40..63 hold
// zero.  A column sits in one 8-lane half-row, so theta's column parity is a
// three-step DPP xor butterfly (quad_perm, quad_perm, row_half_mirror) with
// no LDS traffic; the parities of columns x - 1 and x + 1, and pi with chi's
// two row neighbours, are gathered with ds_bpermute: 10 permutes per round.
// A padding lane takes its column parities from lane 40 (a lane of the
// all-padding group 5, whose parity is zero), does not rotate, and gathers
// pi and chi's three sources from itself: it stays zero with no masking.
struct SpreadLane {
  uint32_t iota, flip, t, cm, cp, s0, s1, s2;  // cm .. s2: source byte addresses for ds_bpermute
  // pi moves A[x', y'] to B[y', 2x' + 3y']: B[x, y] comes from A[3 (y - 3x) mod 5, x]
  static __device__ __forceinline__ uint32_t pi_src(uint32_t x, uint32_t y) {
    return 8 * ((3 * (y + 5 - (3 * x) % 5)) % 5) + x;
  }
  __device__ __forceinline__ explicit SpreadLane(uint32_t lane) {
    const uint32_t x = lane >> 3, y = lane & 7;
    const bool ok = x < 5 && y < 5;
    iota = lane == 0 ? ~0u : 0u;
    // rho as two funnel shifts by t = (32 - r) mod 32 of the halves, swapped
    // first when flip: rotations by 0 and 32 need no other case
    const uint32_t rho = ok ? (uint32_t)kKeccak.rho[x + 5 * y] : 0u, rr = rho & 31;
    flip = (rho >= 32) != (rr == 0) ? ~0u : 0u;
    t = (32 - rr) & 31;
    cm = ok ? 4 * (8 * ((x + 4) % 5) + y) : 4 * 40;
    cp = ok ? 4 * (8 * ((x + 1) % 5) + y) : 4 * 40;
    s0 = ok ? 4 * pi_src(x, y) : 4 * lane;
    s1 = ok ? 4 * pi_src((x + 1) % 5, y) : 4 * lane;
    s2 = ok ? 4 * pi_src((x + 2) % 5, y) : 4 * lane;
  }
  // message word this lane absorbs (x + 5y < 17), or -1
  __device__ __forceinline__ static int word_of(uint32_t lane) {
    const uint32_t x = lane >> 3, y = lane & 7;
    return x < 5 && y < 5 && x + 5 * y < 17 ? (int)(x + 5 * y) : -1;
  }
};

__device__ __forceinline__ uint32_t bperm(uint32_t addr, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_ds_bpermute((int)addr, (int)v);
}
// xor over the lane's 8-lane half-row (all eight lanes get it); every
// source lane lies in the lane's own row, so each step is one xor with a
// DPP operand (the update form lets the compiler fold the move into it)
__device__ __forceinline__ uint32_t xor8(uint32_t v) {
  v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, true);   // quad_perm [1, 0, 3, 2]
  v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, true);   // quad_perm [2, 3, 0, 1]
  v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, true);  // row_half_mirror
  return v;
}

__device__ __forceinline__ void keccak_spread(uint32_t &l, uint32_t &h, const SpreadLane &c) {
#pragma unroll
  for (int ir = 0; ir < 24; ir++) {
    // theta: A ^= C[x - 1] ^ rotl(C[x + 1], 1)
    const uint32_t cl = xor8(l), ch = xor8(h);
    const uint32_t cml = bperm(c.cm, cl), cmh = bperm(c.cm, ch);
    const uint32_t cpl = bperm(c.cp, cl), cph = bperm(c.cp, ch);
    l = xor3(l, cml, __builtin_amdgcn_alignbit(cpl, cph, 31));
    h = xor3(h, cmh, __builtin_amdgcn_alignbit(cph, cpl, 31));
    // rho: rotate this lane's word by its own offset
    const uint32_t a = c.flip ? h : l, b = c.flip ? l : h;
    const uint32_t nl = __builtin_amdgcn_alignbit(a, b, c.t);
    const uint32_t nh = __builtin_amdgcn_alignbit(b, a, c.t);
    // pi + chi: B[x, y] ^ (~B[x + 1, y] & B[x + 2, y])
    l = chi3(bperm(c.s0, nl), bperm(c.s1, nl), bperm(c.s2, nl));
    h = chi3(bperm(c.s0, nh), bperm(c.s1, nh), bperm(c.s2, nh));
    // iota (unrolled: the constants are immediates)
    l ^= (uint32_t)kKeccak.rc[ir] & c.iota;
    h ^= (uint32_t)(kKeccak.rc[ir] >> 32) & c.iota;
  }
}

// Sponge over nb padded blocks, state spread over the wave; every lane must
// call it with the same m and nb.  Digest word q ends on lane 8q (q < 4).
__device__ __forceinline__ void sponge_spread(const uint64_t *m, uint32_t nb, uint32_t lane, const SpreadLane &c,
                                              uint32_t &l, uint32_t &h) {
  l = 0;
  h = 0;
  const int w = SpreadLane::word_of(lane);
  for (uint32_t b = 0; b < nb; b++) {
    if (w >= 0) {
      const uint64_t v = m[17 * b + (uint32_t)w];
      l ^= (uint32_t)v;
      h ^= (uint32_t)(v >> 32);
    }
    keccak_spread(l, h, c);
  }
}
__device__ __forceinline__ bool spread_digest_lane(uint32_t lane) { return (lane & 7) == 0 && lane < 32; }

__global__ __launch_bounds__(256) void k_hc_perm_spread(Pools P, const ref_t *list, const uint64_t *slot,
                                                        const uint32_t *enc_len, uint32_t n,
                                                        const uint64_t *scratch) {
  if (P.cnt->graph_bad) return;  // a replayed pass whose batch is unlike the captured one
  const uint32_t lane = __lane_id();
  const uint32_t i = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  uint32_t enc = i < n ? enc_len[i] : 0;
  uint32_t nb = enc >= 32 ? keccak_blocks(enc) : 0;  // uniform over the wave
  SpreadLane c(lane);
  uint32_t l, h;
  sponge_spread(nb ? scratch + slot[i] / 8 : scratch, nb, lane, c, l, h);
  if (nb == 0) return;
  ref_t r = list[i];
  if (spread_digest_lane(lane)) ((uint64_t *)node_hash(P, r))[lane >> 3] = ((uint64_t)h << 32) | l;
  if (lane == 0) node_done(P, r, 32);
}

// Small inner levels in one launch: one node per wave; lanes 0..15 encode it
// (encode_inner_node), then the wave runs the sponge with the state spread
// over it, reading the message from the node's LDS copy (a generic-encoded
// node's from its scratch slot, which the block barrier makes visible).
__global__ __launch_bounds__(256) void k_hc_small(Pools P, const ref_t *__restrict__ list,
                                                  const uint64_t *__restrict__ slot, uint32_t n,
                                                  uint64_t *__restrict__ scratch, uint32_t *__restrict__ enc_len) {
  if (P.cnt->graph_bad) return;  // a replayed pass whose batch is unlike the captured one
  // latency-bound waves: ahead of a concurrent launch's permutations (the
  // leaf level's, or a level's bulk) on a shared SIMD
  __builtin_amdgcn_s_setprio(2);
  __shared__ uint64_t lds[256 / 64][kInnerWords];
  const uint32_t lane = __lane_id(), g = threadIdx.x / 64;
  const uint32_t i = blockIdx.x * (256 / 64) + g;
  const bool active = i < n;
  const ref_t r = active ? list[i] : REF_NULL;
  uint64_t *out = active ? scratch + slot[i] / 8 : scratch;
  bool in_lds = false;
  uint32_t enc = encode_inner_node(P, r, active, lane < kInnerGroup, lane & (kInnerGroup - 1), lds[g], out, &in_lds);
  enc = (uint32_t)__shfl((int)enc, 0, 64);
  in_lds = __shfl((int)in_lds, 0, 64) != 0;
  if (active && lane == 0) enc_len[i] = enc;
  __syncthreads();  // the node's message words are visible to all its lanes
  const uint32_t nb = active && enc >= 32 ? keccak_blocks(enc) : 0u;  // uniform over the wave
  SpreadLane c(lane);
  uint32_t l, h;
  // a fast node's message stays in LDS (at most 4 blocks: 68 of the 80 words)
  sponge_spread(in_lds ? lds[g] : out, nb, lane, c, l, h);
  if (nb == 0) return;
  if (spread_digest_lane(lane)) ((uint64_t *)node_hash(P, r))[lane >> 3] = ((uint64_t)h << 32) | l;
  if (lane == 0) node_done(P, r, 32);
}

// Two states per wave, one per 32-lane half (mid-size levels: one
// ds_bpermute serves both states, halving the permutes per node).  In a
// half, A[x, y] of columns 0 .. 3 sits in lane 8x + y and column 4's word y
// in lane {5, 6, 7, 13, 14}[y] (the first row's spare lanes); the other
// seven lanes are padding, which may hold anything: they are masked out of
// the column parities and are no lane's source.  Columns 0 .. 3 take their
// parity from an xor butterfly over the masked 8-lane group, column 4 from
// one over the first 16-lane row; lane 5 of a half holds C[4], lane 8x
// C[x], for the neighbour gathers.
struct SpreadLane2 {
  uint32_t m03, m4, cc4, iota, flip, t, cm, cp, s0, s1, s2;  // cm .. s2: ds_bpermute byte addresses
  static __device__ __forceinline__ uint32_t lane_of(uint32_t x, uint32_t y) {
    return x < 4 ? 8 * x + y : (y < 3 ? 5 + y : 10 + y);  // column 4: lanes 5, 6, 7, 13, 14
  }
  // (x, y) of half-lane l, or x = 5 for padding
  static __device__ __forceinline__ void role(uint32_t l, uint32_t &x, uint32_t &y) {
    const uint32_t g = l >> 3, r = l & 7;
    if (g < 4 && r < 5) {
      x = g;
      y = r;
    } else if (l == 5 || l == 6 || l == 7 || l == 13 || l == 14) {
      x = 4;
      y = l < 8 ? l - 5 : l - 10;
    } else {
      x = 5;
      y = 0;
    }
  }
  __device__ __forceinline__ explicit SpreadLane2(uint32_t lane) {
    const uint32_t h = lane >> 5, l = lane & 31, base = 32 * h;
    uint32_t x, y;
    role(l, x, y);
    const bool ok = x < 5;
    m03 = x < 4 ? ~0u : 0u;
    m4 = x == 4 ? ~0u : 0u;
    cc4 = l == 5 ? ~0u : 0u;
    iota = l == 0 ? ~0u : 0u;
    const uint32_t rho = ok ? (uint32_t)kKeccak.rho[x + 5 * y] : 0u, rr = rho & 31;
    flip = (rho >= 32) != (rr == 0) ? ~0u : 0u;
    t = (32 - rr) & 31;
    if (ok) {
      const uint32_t xm = (x + 4) % 5, xp = (x + 1) % 5;
      cm = 4 * (base + (xm == 4 ? 5u : 8 * xm));
      cp = 4 * (base + (xp == 4 ? 5u : 8 * xp));
      // pi: B[x, y] comes from A[3 (y - 3x) mod 5, x]
      s0 = 4 * (base + lane_of((3 * (y + 5 - (3 * x) % 5)) % 5, x));
      s1 = 4 * (base + lane_of((3 * (y + 5 - (3 * ((x + 1) % 5)) % 5)) % 5, (x + 1) % 5));
      s2 = 4 * (base + lane_of((3 * (y + 5 - (3 * ((x + 2) % 5)) % 5)) % 5, (x + 2) % 5));
    } else {
      cm = cp = s0 = s1 = s2 = 4 * lane;
    }
  }
  // message word this lane absorbs (x + 5y < 17), or -1
  __device__ __forceinline__ static int word_of(uint32_t lane) {
    uint32_t x, y;
    role(lane & 31, x, y);
    return x < 5 && x + 5 * y < 17 ? (int)(x + 5 * y) : -1;
  }
};

// xor over the lane's 16-lane row
__device__ __forceinline__ uint32_t xor16(uint32_t v) {
  v = xor8(v);
  return v ^ (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, true);  // row_mirror
}

__device__ __forceinline__ void keccak_spread2(uint32_t &l, uint32_t &h, const SpreadLane2 &c) {
#pragma unroll
  for (int ir = 0; ir < 24; ir++) {
    // theta's column parities: C[x] in every lane of group x, C[4] in the
    // first row; the holders' word for the gathers below
    const uint32_t c8l = xor8(l & c.m03), c8h = xor8(h & c.m03);
    const uint32_t c16l = xor16(l & c.m4), c16h = xor16(h & c.m4);
    const uint32_t ccl = (c16l & c.cc4) | (c8l & ~c.cc4), cch = (c16h & c.cc4) | (c8h & ~c.cc4);
    const uint32_t cml = bperm(c.cm, ccl), cmh = bperm(c.cm, cch);
    const uint32_t cpl = bperm(c.cp, ccl), cph = bperm(c.cp, cch);
    l = xor3(l, cml, __builtin_amdgcn_alignbit(cpl, cph, 31));
    h = xor3(h, cmh, __builtin_amdgcn_alignbit(cph, cpl, 31));
    // rho
    const uint32_t a = c.flip ? h : l, b = c.flip ? l : h;
    const uint32_t nl = __builtin_amdgcn_alignbit(a, b, c.t);
    const uint32_t nh = __builtin_amdgcn_alignbit(b, a, c.t);
    // pi + chi
    l = chi3(bperm(c.s0, nl), bperm(c.s1, nl), bperm(c.s2, nl));
    h = chi3(bperm(c.s0, nh), bperm(c.s1, nh), bperm(c.s2, nh));
    // iota
    l ^= (uint32_t)kKeccak.rc[ir] & c.iota;
    h ^= (uint32_t)(kKeccak.rc[ir] >> 32) & c.iota;
  }
}

// Two nodes in one wave, one per half: nodes 2 * pair and 2 * pair + 1 of
// list[0 .. n).  Lanes 0 .. 15 of a half encode its node (encode_inner_node,
// the message in lds[half] when it fits), then both halves run the sponge
// together, each over its own message (the wave runs as many blocks as its
// taller node; a half keeps its digest after its own last block).  Every
// thread of the block calls it (it holds a barrier).
__device__ __forceinline__ void hash_pair_spread2(const Pools &P, const ref_t *__restrict__ list,
                                                  const uint64_t *__restrict__ slot, uint32_t n,
                                                  uint64_t *__restrict__ scratch, uint32_t *__restrict__ enc_len,
                                                  uint32_t pair, uint64_t (*lds)[kInnerWords]) {
  const uint32_t lane = __lane_id(), half = lane >> 5, hl = lane & 31;
  const uint32_t i = 2 * pair + half;
  const bool active = i < n;
  const ref_t r = active ? list[i] : REF_NULL;
  uint64_t *out = active ? scratch + slot[i] / 8 : scratch;
  bool in_lds = false;
  uint32_t enc = encode_inner_node(P, r, active, hl < kInnerGroup, hl & (kInnerGroup - 1), lds[half], out, &in_lds);
  enc = (uint32_t)__shfl((int)enc, 32 * half, 64);
  in_lds = __shfl((int)in_lds, 32 * half, 64) != 0;
  if (active && hl == 0) enc_len[i] = enc;
  __syncthreads();  // the nodes' message words are visible to all their lanes
  const uint32_t nb = active && enc >= 32 ? keccak_blocks(enc) : 0u;  // uniform over the half
  const uint32_t nbmax = max((uint32_t)__shfl((int)nb, 0, 64), (uint32_t)__shfl((int)nb, 32, 64));
  const SpreadLane2 c(lane);
  const int w = SpreadLane2::word_of(lane);
  const uint64_t *m = in_lds ? lds[half] : out;
  uint32_t l = 0, h = 0, dl = 0, dh = 0;
  for (uint32_t b = 0; b < nbmax; b++) {
    if (w >= 0 && b < nb) {
      const uint64_t v = m[17 * b + (uint32_t)w];
      l ^= (uint32_t)v;
      h ^= (uint32_t)(v >> 32);
    }
    keccak_spread2(l, h, c);
    if (b + 1 == nb) {
      dl = l;
      dh = h;
    }
  }
  if (nb == 0) return;
  // digest word q is A[q, 0]: half-lane 8q
  if ((hl & 7) == 0 && hl < 32) ((uint64_t *)node_hash(P, r))[hl >> 3] = ((uint64_t)dh << 32) | dl;
  if (hl == 0) node_done(P, r, 32);
}

// Small inner levels of many nodes: two nodes per wave (hash_pair_spread2).
__global__ __launch_bounds__(256) void k_hc_small2(Pools P, const ref_t *__restrict__ list,
                                                   const uint64_t *__restrict__ slot, uint32_t n,
                                                   uint64_t *__restrict__ scratch, uint32_t *__restrict__ enc_len) {
  if (P.cnt->graph_bad) return;  // a replayed pass whose batch is unlike the captured one
  __builtin_amdgcn_s_setprio(2);  // as in k_hc_small
  __shared__ uint64_t lds[256 / 32][kInnerWords];
  const uint32_t wave = threadIdx.x / 64;
  hash_pair_spread2(P, list, slot, n, scratch, enc_len, 4 * blockIdx.x + wave, lds + 2 * wave);
}

__global__ __launch_bounds__(256) void k_hc_roots(Pools P, uint32_t ntries, const unsigned long long *parts) {
  uint32_t t = grid_tid();
  // the root's hash first: its two dependent loads overlap the counts' reduction
  uint32_t hs = 0;
  uint4 h[2] = {make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0)};
  if (t < ntries) {
    const ref_t r = P.tries[t].root;
    if (ref_type(r) == T_SHORT || ref_type(r) == T_BRANCH) {
      hs = node_hash_size(P, r);
      const uint4 *src = (const uint4 *)node_hash(P, r);  // offset 0 of a 64- or 128-byte node
      h[0] = src[0];
      h[1] = src[1];
    }
  }
  if (t == 0) P.cnt->n_dirty = 0;  // every listed node has been hashed
  if (parts && blockIdx.x == 0) {  // the pass's work counts (uniform branch)
    static_assert(kHcCountParts == 256, "one partial per thread");
    __shared__ unsigned long long sh[4][4];
    unsigned long long v[4];
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = parts[4 * threadIdx.x + q];
#pragma unroll
    for (int q = 0; q < 4; q++)
      for (int off = 32; off > 0; off >>= 1) v[q] += (unsigned long long)__shfl_xor((long long)v[q], off);
    if (__lane_id() == 0)
#pragma unroll
      for (int q = 0; q < 4; q++) sh[threadIdx.x >> 6][q] = v[q];
    __syncthreads();
    if (threadIdx.x < 4) {
      unsigned long long *dst = threadIdx.x == 0 ? &P.cnt->n_perms
                                : threadIdx.x == 1 ? &P.cnt->enc_bytes
                                : threadIdx.x == 2 ? &P.cnt->leaf_perms
                                                   : &P.cnt->leaf_bytes;
      *dst += sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
    }
  }
  if (t >= ntries) return;
  // bytes past the hash size (a root encoding under 32 bytes) read as zero
  uint32_t w[8] = {h[0].x, h[0].y, h[0].z, h[0].w, h[1].x, h[1].y, h[1].z, h[1].w};
#pragma unroll
  for (uint32_t q = 0; q < 8; q++) {
    const uint32_t keep = hs >= 4 * q + 4 ? 4u : hs > 4 * q ? hs - 4 * q : 0u;
    w[q] = keep == 4 ? w[q] : keep ? w[q] & ((1u << (8 * keep)) - 1) : 0u;
  }
  TrieHdr &T = P.tries[t];
  uint64_t *dst = (uint64_t *)T.root_hash;  // offset 16 of the header
#pragma unroll
  for (uint32_t q = 0; q < 4; q++) dst[q] = ((uint64_t)w[2 * q + 1] << 32) | w[2 * q];
  T.root_hash_size = hs;
  if (t == 0) {
    uint32_t *d0 = (uint32_t *)P.cnt->root0;
#pragma unroll
    for (uint32_t q = 0; q < 8; q++) d0[q] = w[q];
    P.cnt->root0_size = hs;
  }
}

// ---------------------------------------------- standalone Keccak (test hook) --

__global__ __launch_bounds__(256) void k_kb_slots(const uint32_t *len, uint32_t n, uint64_t *slot) {
  uint32_t i = grid_tid();
  if (i < n) slot[i] = 136ull * keccak_blocks(len[i]);
}

__global__ __launch_bounds__(256) void k_kb_pad(const uint8_t *data, const uint64_t *off, const uint32_t *len,
                                                uint32_t n, const uint64_t *slot, uint64_t *scratch) {
  uint32_t i = grid_tid();
  if (i >= n) return;
  uint32_t L = len[i];
  WordWriter w(scratch + slot[i] / 8);
  if (L) w.mem_bytes(data + off[i], L);
  w.finish_keccak(keccak_blocks(L) * 17);
}

__global__ __launch_bounds__(256) void k_kb_lane(const uint64_t *slot, const uint32_t *len, uint32_t n,
                                                 const uint64_t *scratch, uint8_t *out) {
  uint32_t i = grid_tid();
  if (i >= n) return;
  uint64_t s[25];
  keccak_lane(scratch + slot[i] / 8, keccak_blocks(len[i]), s);
  for (int q = 0; q < 4; q++) ((uint64_t *)(out + 32 * (size_t)i))[q] = s[q];
}

__global__ __launch_bounds__(256) void k_kb_spread(const uint64_t *slot, const uint32_t *len, uint32_t n,
                                                   const uint64_t *scratch, uint8_t *out) {
  const uint32_t lane = __lane_id();
  const uint32_t i = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  uint32_t nb = i < n ? keccak_blocks(len[i]) : 0;
  SpreadLane c(lane);
  uint32_t l, h;
  sponge_spread(nb ? scratch + slot[i] / 8 : scratch, nb, lane, c, l, h);
  if (nb && spread_digest_lane(lane)) ((uint64_t *)(out + 32 * (size_t)i))[lane >> 3] = ((uint64_t)h << 32) | l;
}

// ------------------------------------------------------------- node export --

// Byte-granular writer for the export (encodings are concatenated unaligned).
struct ByteWriter {
  uint8_t *out;
  uint32_t pos = 0;
  __device__ __forceinline__ explicit ByteWriter(uint8_t *o) : out(o) {}
  __device__ __forceinline__ void inline_bytes(uint64_t b, uint32_t len) {
    for (uint32_t q = 0; q < len; q++) out[pos++] = (uint8_t)(b >> (8 * q));
  }
  __device__ __forceinline__ void mem_bytes(const uint8_t *src, uint32_t len) {
    for (uint32_t q = 0; q < len; q++) out[pos++] = src[q];
  }
  __device__ __forceinline__ void rlp_mem(const uint8_t *src, uint32_t len) {
    if (len == 1 && src[0] <= 0x7f) {
      inline_bytes(src[0], 1);
      return;
    }
    uint32_t hn;
    uint64_t h = rlp_hdr(len, 0x80, 0xB7, &hn);
    inline_bytes(h, hn);
    mem_bytes(src, len);
  }
  __device__ __forceinline__ void rlp_compact(const uint8_t *key, uint32_t s, uint32_t cnt, uint32_t term) {
    uint32_t kc = cnt / 2 + 1;
    uint32_t flag = (term << 5) | ((cnt & 1) << 4) | ((cnt & 1) ? packed_nibble(key, s) : 0u);
    if (kc == 1) {
      inline_bytes(flag, 1);
      return;
    }
    uint32_t hn;
    uint64_t h = rlp_hdr(kc, 0x80, 0xB7, &hn);
    inline_bytes(h, hn);
    inline_bytes(flag, 1);
    uint32_t first = s + (cnt & 1);
    for (uint32_t q = 0; q + 1 < kc; q++)
      out[pos++] = (uint8_t)((packed_nibble(key, first + 2 * q) << 4) | packed_nibble(key, first + 2 * q + 1));
  }
};

__global__ __launch_bounds__(256) void k_enc_len(Pools P, const ref_t *refs, uint32_t m, uint64_t *len) {
  uint32_t i = grid_tid();
  if (i >= m) return;
  uint32_t pay = node_payload(P, refs[i]);
  len[i] = rlp_hdr_len(pay) + pay;
}

// Node export: RLP encoding at offs[i] and the 32-byte hash slot of node i.
__global__ __launch_bounds__(256) void k_enc_write(Pools P, const ref_t *refs, uint32_t m, const uint64_t *offs,
                                                   uint8_t *encs, uint8_t *hashes) {
  uint32_t i = grid_tid();
  if (i >= m) return;
  ref_t r = refs[i];
  uint32_t pay = node_payload(P, r);
  ByteWriter w(encs + offs[i]);
  emit_node(P, r, pay, w);
  const uint8_t *h = node_hash(P, r);
  uint32_t hs = node_hash_size(P, r);
  for (int q = 0; q < 32; q++) hashes[32 * (size_t)i + q] = q < (int)hs ? h[q] : 0;
}

}  // namespace

size_t hc_temp_bytes(uint32_t n) { return scan_temp_bytes(n > kHcBins ? n : kHcBins); }

hipError_t hc_classify(const Pools &P, const ref_t *list, uint32_t n, HcMeta *meta, uint32_t *cursor, ref_t *out,
                       uint64_t *slot, uint32_t *packed, void *temp, hipStream_t s, uint64_t *lmeta_tmp,
                       uint64_t *lmeta, LeafRecords rec) {
  // the bins, and the pass's work counters (permutations, bytes, leaf
  // permutations, leaf bytes)
  static_assert(kHcBins >= 8, "one launch clears both");
  k_hc_zero<<<nblk(kHcBins), kB, 0, s>>>(meta->hist, kHcBins, (uint32_t *)&P.cnt->n_perms, 8);
  GMPT_LAUNCH_CHECK();
  if (n == 0) {
    k_hc_zero<<<1, kB, 0, s>>>(&meta->slot_lo, 4);  // slot_lo, nvalid, scratch_bytes
    return hipGetLastError();
  }
  unsigned tiles = (unsigned)((n + kHcTile - 1) / kHcTile);
  k_hc_hist<<<tiles, kB, 0, s>>>(P, list, n, meta->hist, packed, lmeta ? lmeta_tmp : nullptr, rec);
  GMPT_LAUNCH_CHECK();
  hipError_t e = scan_exclusive_u32(meta->hist, cursor, kHcBins, &meta->nvalid, temp, s);
  if (e != hipSuccess) return e;
  k_hc_scatter<<<tiles, kB, 0, s>>>(P, list, packed, n, cursor, out, slot, lmeta_tmp, lmeta, rec, meta);
  GMPT_LAUNCH_CHECK();
  // slot offsets of the sorted nodes that have a slot: [slot_lo, nvalid)
  return scan_exclusive_u64(slot, slot, n, &meta->scratch_bytes, temp, s, nullptr, &meta->slot_lo);
}

hipError_t hc_hash_range(const Pools &P, const ref_t *list, const uint64_t *slot, uint32_t n, uint64_t *scratch,
                         uint32_t *enc_len, bool leaves, uint32_t spread_max, hipStream_t s, hipEvent_t perm_begin,
                         hipEvent_t perm_end, uint32_t n_ext, const uint32_t *cls) {
  if (n == 0) return hipSuccess;
  if (!leaves && n <= spread_max) {  // small inner level: encode and hash in one launch
    if (perm_begin) {
      hipError_t e = hipEventRecord(perm_begin, s);
      if (e != hipSuccess) return e;
    }
    if (n >= kSpread2Min)  // enough nodes that the permutes' throughput, not one node's latency, sets the time
      k_hc_small2<<<(n + 7) / 8, kB, 0, s>>>(P, list, slot, n, scratch, enc_len);
    else
      k_hc_small<<<(n + 3) / 4, kB, 0, s>>>(P, list, slot, n, scratch, enc_len);
    GMPT_LAUNCH_CHECK();
    if (perm_end) return hipEventRecord(perm_end, s);
    return hipSuccess;
  }
  // a level of branches of at most 4 blocks and extensions (all but a
  // sixteenth of its nodes): one launch with the messages in LDS, the rest
  // generic inside it; rows as tall as the level's tallest class needs (1, 2
  // or 4 blocks: taller rows leave fewer waves per CU)
  const uint32_t c01 = cls ? cls[0] + cls[1] : 0u, c23 = cls ? cls[2] + cls[3] : 0u;
  const uint32_t cext = cls ? cls[kExtClass] : 0u;
  const int rows_nb = !cls || leaves ? 0
                      : 16ull * ((uint64_t)c01 + cext) >= 15ull * n       ? (cls[1] ? 2 : 1)
                      : 16ull * ((uint64_t)c01 + c23 + cext) >= 15ull * n ? 4
                                                                            : 0;
  if (rows_nb) {
    // the few nodes taller than the rows (classes rows_nb .. 6, between the
    // row classes and the extensions) would take the generic encoder and hold
    // the launch open long after the bulk: the launch's first blocks, which
    // are dispatched first, encode and hash them concurrently with the bulk
    const uint32_t cov = std::min(n, rows_nb == 4 ? c01 + c23 : rows_nb == 2 ? c01 : cls[0]);
    const uint32_t multi = n - std::min(n, cov + cext);
    const uint32_t lo = multi ? cov : 0u, hi = lo + multi;
    const uint32_t nx = (multi + 1) / 2;  // the launch's first blocks: two of them per block
    const unsigned grid = (unsigned)((n + 63) / 64 + nx);
    if (rows_nb == 4)
      k_hc_inner_wave<4><<<grid, 64, 0, s>>>(P, list, slot, n, scratch, enc_len, lo, hi, nx);
    else if (rows_nb == 2)
      k_hc_inner_wave<2><<<grid, 64, 0, s>>>(P, list, slot, n, scratch, enc_len, lo, hi, nx);
    else
      k_hc_inner_wave<1><<<grid, 64, 0, s>>>(P, list, slot, n, scratch, enc_len, lo, hi, nx);
    GMPT_LAUNCH_CHECK();
    return hipSuccess;
  }
  // the level's extension class, sorted last: its own launch when it is large
  // enough to fill the GPU (a few thousand ride along in the branch kernels
  // at no cost, but would pay a launch's latency)
  if (!leaves && n_ext >= kExtDirectMin) {
    const uint32_t m = n - n_ext;
    k_hc_ext_direct<<<nblk(n_ext), kB, 0, s>>>(P, list + m, slot + m, n_ext, scratch, enc_len + m);
    GMPT_LAUNCH_CHECK();
    n = m;
    if (n == 0) return hipSuccess;
  }
  if (leaves)
    k_hc_encode_leaf<<<(n + kLeafNodes - 1) / kLeafNodes, kB, 0, s>>>(P, list, slot, n, scratch, enc_len);
  else
    k_hc_encode_inner<<<(n + kInnerNodes - 1) / kInnerNodes, kB, 0, s>>>(P, list, slot, n, scratch, enc_len);
  GMPT_LAUNCH_CHECK();
  if (perm_begin) {
    hipError_t e = hipEventRecord(perm_begin, s);
    if (e != hipSuccess) return e;
  }
  if (n <= spread_max)
    k_hc_perm_spread<<<(n + 3) / 4, kB, 0, s>>>(P, list, slot, enc_len, n, scratch);
  else
    k_hc_perm_lane<<<nblk(n), kB, 0, s>>>(P, list, slot, enc_len, n, scratch);
  GMPT_LAUNCH_CHECK();
  if (perm_end) return hipEventRecord(perm_end, s);
  return hipSuccess;
}

hipError_t hc_hash_leaves_direct(const Pools &P, const ref_t *list, const uint64_t *lmeta, const HcMeta *meta,
                                 uint32_t n_upper, uint32_t *enc_len, hipStream_t s) {
  if (n_upper == 0) return hipSuccess;
  k_hc_leaf_direct<<<nblk(n_upper), kB, 0, s>>>(P, list, lmeta, meta->hist, enc_len);
  return hipGetLastError();
}

hipError_t hc_tail_leaves(const Pools &P, const ref_t *list, uint32_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_hc_tail_leaves<<<(unsigned)((16ull * n + 255) / 256), kB, 0, s>>>(P, list, n);
  return hipGetLastError();
}

hipError_t hc_count(const uint32_t *enc_len, const HcMeta *meta, uint32_t n, uint32_t n_leaf,
                    unsigned long long *parts, hipStream_t s) {
  k_hc_count<<<kHcCountParts, kB, 0, s>>>(parts, enc_len, &meta->nvalid, n_leaf);
  return hipGetLastError();
}

namespace {
__global__ __launch_bounds__(256) void k_clear(Counters *cnt, TrieHdr *tries, uint32_t ntries) {
  uint32_t t = grid_tid();
  if (t < sizeof(Counters) / 4) ((uint32_t *)cnt)[t] = 0;
  for (uint32_t q = t; q < ntries * (uint32_t)(sizeof(TrieHdr) / 4); q += gridDim.x * blockDim.x)
    ((uint32_t *)tries)[q] = 0;
}
}  // namespace

namespace {
// A replayed pass is valid when the batch sorted without fallback, raised no
// error, and produced the captured pass's dirty-list length, element count
// and hash bins; otherwise it marks the pass bad (the hash kernels, which
// launch with the captured level ranges, then skip).
__global__ __launch_bounds__(256) void k_graph_check(Pools P, const uint32_t *__restrict__ hist,
                                                     const uint32_t *__restrict__ plan, uint32_t nbins,
                                                     uint32_t n_dirty, uint32_t ne, const uint32_t *scal) {
  __shared__ uint32_t bad;
  if (threadIdx.x == 0)
    bad = P.cnt->err != 0 || P.cnt->flag_fallback != 0 || P.cnt->n_dirty != n_dirty || scal[SC_NE_INDEX] != ne;
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < nbins; b += blockDim.x)
    if (hist[b] != plan[b]) bad = 1;
  __syncthreads();
  if (threadIdx.x == 0 && bad) P.cnt->graph_bad = 1;
}
}  // namespace

hipError_t hc_graph_check(const Pools &P, const HcMeta *meta, const uint32_t *plan_hist, uint32_t n_dirty, uint32_t ne,
                          const uint32_t *scal, hipStream_t s) {
  k_graph_check<<<1, kB, 0, s>>>(P, meta->hist, plan_hist, kHcBins, n_dirty, ne, scal);
  return hipGetLastError();
}

hipError_t hc_clear(Counters *cnt, TrieHdr *tries, uint32_t ntries, hipStream_t s) {
  size_t words = std::max<size_t>(sizeof(Counters) / 4, (size_t)ntries * sizeof(TrieHdr) / 4);
  k_clear<<<(unsigned)std::min<size_t>(nblk(words), 1024), kB, 0, s>>>(cnt, tries, ntries);
  return hipGetLastError();
}

hipError_t hc_roots(const Pools &P, uint32_t ntries, const unsigned long long *parts, hipStream_t s) {
  k_hc_roots<<<nblk(ntries), kB, 0, s>>>(P, ntries, parts);
  return hipGetLastError();
}

hipError_t hc_enc_len(const Pools &P, const ref_t *refs, uint32_t m, uint64_t *len, hipStream_t s) {
  if (m == 0) return hipSuccess;
  k_enc_len<<<nblk(m), kB, 0, s>>>(P, refs, m, len);
  return hipGetLastError();
}

hipError_t hc_enc_write(const Pools &P, const ref_t *refs, uint32_t m, const uint64_t *offs, uint8_t *encs,
                        uint8_t *hashes, hipStream_t s) {
  if (m == 0) return hipSuccess;
  k_enc_write<<<nblk(m), kB, 0, s>>>(P, refs, m, offs, encs, hashes);
  return hipGetLastError();
}

uint64_t keccak_scratch_bytes(const uint32_t *len, uint32_t n) {
  uint64_t t = 0;
  for (uint32_t i = 0; i < n; i++) t += 136ull * (len[i] / 136 + 1);
  return t;
}

hipError_t keccak_batch(const uint8_t *data, const uint64_t *off, const uint32_t *len, uint32_t n, uint8_t *out,
                        uint64_t *slot, uint64_t *scratch, void *temp, int spread, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_kb_slots<<<nblk(n), kB, 0, s>>>(len, n, slot);
  GMPT_LAUNCH_CHECK();
  hipError_t e = scan_exclusive_u64(slot, slot, n, nullptr, temp, s);
  if (e != hipSuccess) return e;
  k_kb_pad<<<nblk(n), kB, 0, s>>>(data, off, len, n, slot, scratch);
  GMPT_LAUNCH_CHECK();
  if (spread)
    k_kb_spread<<<(n + 3) / 4, kB, 0, s>>>(slot, len, n, scratch, out);
  else
    k_kb_lane<<<nblk(n), kB, 0, s>>>(slot, len, n, scratch, out);
  return hipGetLastError();
}

}  // namespace gmpt
