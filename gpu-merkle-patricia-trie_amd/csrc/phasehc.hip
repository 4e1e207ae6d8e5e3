// phasehc.hip -- PhaseHC: bottom-up one-pass hashing of the dirty nodes.
//
// Reference: hash_onepass_v2 (mpt-with-compress/include/mpt/gpu_mpt.cuh:1162-1191)
// counts children per node (mark phase, gpu_mpt_kernels.cuh:2696-2728) and
// lets the last-arriving warp encode and hash each node (update phase,
// :2885-3241).  Here every dirty node carries its nibble depth, so the hash
// order is fixed without counters: all leaf short nodes first (they depend on
// nothing), then non-leaf nodes by decreasing depth (a child is always deeper
// than its parent).  Inside a level, nodes are grouped by their Keccak block
// count so a wave's lanes run the same number of permutations.
//
// Node encoding follows the reference's encoders exactly
// (FullNode::encode node.cuh:414-452, ShortNode::encode node.cuh:526-559,
// rlp.cuh:118-155): every child is written with write_bytes, i.e. a child
// whose encoding is shorter than 32 bytes is string-wrapped (not inlined),
// a value node contributes its raw bytes, and a node whose own encoding is
// shorter than 32 bytes keeps the raw encoding instead of a digest.
#include "phasehc.h"
#include "keccak.h"
#include "prims.h"

namespace gmpt {

namespace {

constexpr int kB = 256;
constexpr uint32_t kHcTile = 2048;  // dirty nodes per binning block

__device__ __forceinline__ uint32_t grid_tid() { return blockIdx.x * blockDim.x + threadIdx.x; }
inline unsigned nblk(size_t n) { return (unsigned)((n + kB - 1) / kB); }

__device__ __forceinline__ uint32_t be_bytes(uint64_t v) {
  uint32_t n = 1;
  while (v >>= 8) n++;
  return n;
}
// RLP header length for a string/list payload of L bytes.
__device__ __forceinline__ uint32_t rlp_hdr_len(uint32_t L) { return L < 56 ? 1u : 1u + be_bytes(L); }

// write_bytes_size (rlp.cuh:36-51); first = the first byte (only used when L == 1)
__device__ __forceinline__ uint32_t rlp_item_len(uint32_t L, uint32_t first) {
  if (L == 1 && first <= 0x7f) return 1;
  return rlp_hdr_len(L) + L;
}

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

// Number of Keccak blocks a node's message needs, estimated before its
// children are hashed (children assumed to carry 32-byte digests).
__device__ uint32_t estimate_blocks(const Pools &P, ref_t r, bool *leaf) {
  uint32_t t = ref_type(r), i = ref_idx(r);
  uint32_t len;
  *leaf = false;
  if (t == T_SHORT) {
    const Short &S = P.shrt[i];
    uint32_t kc = (S.key_len - S.term) / 2 + 1;
    uint32_t pay = rlp_item_len(kc, 0);
    if (ref_type(S.child) == T_VALUE) {
      *leaf = true;
      pay += rlp_item_len(P.value[ref_idx(S.child)].size, 0x80);
    } else {
      pay += 33;
    }
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
  return len / 136 + 1;
}

__device__ __forceinline__ uint32_t bin_of(const Pools &P, ref_t r) {
  bool leaf;
  uint32_t blocks = estimate_blocks(P, r, &leaf);
  uint32_t cls = blocks > kHcClasses ? kHcClasses - 1 : blocks - 1;
  uint32_t level;
  if (leaf) {
    level = 0;
  } else {
    uint32_t depth = ref_type(r) == T_SHORT ? P.shrt[ref_idx(r)].depth : P.branch[ref_idx(r)].depth;
    level = 1 + (kHcMaxDepth - min(depth, (uint32_t)kHcMaxDepth));
  }
  return level * kHcClasses + cls;
}

// Counting sort of the dirty list by bin.  Counts are aggregated per block in
// LDS so each block issues one global atomic per bin it touches (a global
// atomic per node serialises on the few hot bins).
__global__ __launch_bounds__(256) void k_hc_hist(Pools P, const ref_t *list, uint32_t n, uint32_t *hist) {
  __shared__ uint32_t h[kHcBins];
  for (uint32_t b = threadIdx.x; b < kHcBins; b += blockDim.x) h[b] = 0;
  __syncthreads();
  for (uint32_t i = blockIdx.x * kHcTile + threadIdx.x; i < min(n, (blockIdx.x + 1) * kHcTile); i += blockDim.x) {
    ref_t r = list[i];
    if (r != REF_NULL) atomicAdd(&h[bin_of(P, r)], 1u);  // builds leave null holes
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < kHcBins; b += blockDim.x)
    if (h[b]) atomicAdd(&hist[b], h[b]);
}

__global__ __launch_bounds__(256) void k_hc_scatter(Pools P, const ref_t *list, uint32_t n, uint32_t *cursor,
                                                    ref_t *out) {
  __shared__ uint32_t cnt[kHcBins];
  __shared__ uint32_t base[kHcBins];
  for (uint32_t b = threadIdx.x; b < kHcBins; b += blockDim.x) cnt[b] = 0;
  __syncthreads();
  constexpr int kPer = kHcTile / 256;
  uint32_t bin[kPer], rank[kPer];
  ref_t r[kPer];
#pragma unroll
  for (int q = 0; q < kPer; q++) {
    uint32_t i = blockIdx.x * kHcTile + q * 256 + threadIdx.x;
    r[q] = i < n ? list[i] : REF_NULL;
    if (r[q] != REF_NULL) {
      bin[q] = bin_of(P, r[q]);
      rank[q] = atomicAdd(&cnt[bin[q]], 1u);
    }
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < kHcBins; b += blockDim.x)
    if (cnt[b]) base[b] = atomicAdd(&cursor[b], cnt[b]);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kPer; q++)
    if (r[q] != REF_NULL) out[base[bin[q]] + rank[q]] = r[q];
}

// ------------------------------------------------------------------ hashing --

// One lane's staging block in LDS: message byte k of the current block sits
// at word k/8 of the lane; words are interleaved across the wave's 64 lanes so
// the absorb reads (ds_read_b64) are bank-conflict free.
struct LdsSink {
  uint8_t *lane;   // wave region + lane * 8
  __device__ __forceinline__ void put(uint32_t k, uint32_t v) { lane[(k >> 3) * 512 + (k & 7)] = (uint8_t)v; }
};
// Whole encoding written to global memory (node export).
struct MemSink {
  uint8_t *out;
  __device__ __forceinline__ void put(uint32_t k, uint32_t v) { out[k] = (uint8_t)v; }
};

// Emits message bytes [lo, lo + W) of a node encoding into a sink.
template <typename Sink, uint32_t W>
struct Stage {
  Sink sink;
  uint32_t lo;     // message offset of this block
  uint32_t pos;    // message offset of the next part to emit
  __device__ __forceinline__ void put(uint32_t m, uint32_t v) { sink.put(m - lo, v); }
  __device__ __forceinline__ bool overlaps(uint32_t len) const { return pos < lo + W && pos + len > lo; }
  __device__ __forceinline__ void inline_bytes(uint64_t b, uint32_t len) {
    if (overlaps(len)) {
      for (uint32_t q = 0; q < len; q++) {
        uint32_t m = pos + q;
        if (m >= lo && m < lo + W) put(m, (uint32_t)(b >> (8 * q)) & 255u);
      }
    }
    pos += len;
  }
  __device__ __forceinline__ void mem_bytes(const uint8_t *src, uint32_t len) {
    if (overlaps(len)) {
      uint32_t a = pos > lo ? pos : lo;
      uint32_t b = min(pos + len, lo + W);
      for (uint32_t m = a; m < b; m++) put(m, src[m - pos]);
    }
    pos += len;
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
  // write_bytes of the hex-prefix (compact) form of nibbles [s, s+cnt) of a
  // packed key, with the leaf flag (utils.cuh:372-396)
  __device__ __forceinline__ void rlp_compact(const uint8_t *key, uint32_t s, uint32_t cnt, uint32_t term) {
    uint32_t kc = cnt / 2 + 1;
    uint32_t flag = (term << 5) | ((cnt & 1) << 4) | ((cnt & 1) ? packed_nibble(key, s) : 0u);
    if (kc == 1) {  // single byte <= 0x3f: written raw
      inline_bytes(flag, 1);
      return;
    }
    uint32_t hn;
    uint64_t h = rlp_hdr(kc, 0x80, 0xB7, &hn);
    inline_bytes(h, hn);
    inline_bytes(flag, 1);
    uint32_t first = s + (cnt & 1);
    uint32_t len = kc - 1;
    if (overlaps(len)) {
      uint32_t a = pos > lo ? pos : lo;
      uint32_t b = min(pos + len, lo + W);
      for (uint32_t m = a; m < b; m++) {
        uint32_t q = first + 2 * (m - pos);
        put(m, (packed_nibble(key, q) << 4) | packed_nibble(key, q + 1));
      }
    }
    pos += len;
  }
};

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

// One Keccak block of a node's message, assembled 8 bytes at a time: parts
// are emitted in message order, bytes outside [lo, lo + 136) are skipped and
// completed 64-bit words go to the lane's LDS words (word w at lane + 512 w).
struct Blk {
  uint8_t *lane;
  uint32_t lo;
  uint32_t pos = 0;   // message offset of the next part
  uint64_t acc = 0;   // pending bytes
  uint32_t accn = 0;
  uint32_t nw = 0;    // words written
  __device__ __forceinline__ Blk(uint8_t *l, uint32_t block_lo) : lane(l), lo(block_lo) {}
  __device__ __forceinline__ void push(uint64_t v, uint32_t n) {  // n in 1..8, v's upper bytes zero
    acc |= v << (8 * accn);
    uint32_t tot = accn + n;
    if (tot >= 8) {
      *(uint64_t *)(lane + nw * 512) = acc;
      nw++;
      uint32_t used = 8 - accn;
      acc = used >= 8 ? 0ull : (v >> (8 * used));
      accn = tot - 8;
    } else {
      accn = tot;
    }
  }
  __device__ __forceinline__ void window(uint32_t n, uint32_t *a, uint32_t *e) const {
    *a = pos > lo ? pos : lo;
    *e = min(pos + n, lo + 136);
  }
  __device__ __forceinline__ void inline_bytes(uint64_t b, uint32_t len) {  // len <= 8
    uint32_t a, e;
    window(len, &a, &e);
    if (a < e) {
      uint32_t c = e - a;
      uint64_t v = b >> (8 * (a - pos));
      if (c < 8) v &= (1ull << (8 * c)) - 1;
      push(v, c);
    }
    pos += len;
  }
  __device__ __forceinline__ void mem_bytes(const uint8_t *src, uint32_t len) {
    uint32_t a, e;
    window(len, &a, &e);
    if (a < e) {
      const uint8_t *sp = src + (a - pos);
      uint32_t c = e - a;
      for (; c >= 8; c -= 8, sp += 8) push(load8(sp), 8);
      if (c) push(load_part(sp, c), c);
    }
    pos += len;
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
  // write_bytes of the hex-prefix form of nibbles [s, s + cnt) of a packed
  // key with the leaf flag (utils.cuh:372-396)
  __device__ __forceinline__ void rlp_compact(const uint8_t *key, uint32_t s, uint32_t cnt, uint32_t term) {
    uint32_t kc = cnt / 2 + 1;
    uint32_t flag = (term << 5) | ((cnt & 1) << 4) | ((cnt & 1) ? packed_nibble(key, s) : 0u);
    if (kc == 1) {  // single byte <= 0x3f: written raw
      inline_bytes(flag, 1);
      return;
    }
    uint32_t hn;
    uint64_t h = rlp_hdr(kc, 0x80, 0xB7, &hn);
    inline_bytes(h, hn);
    inline_bytes(flag, 1);
    uint32_t first = s + (cnt & 1);
    if ((first & 1) == 0) {
      mem_bytes(key + first / 2, kc - 1);
      return;
    }
    // nibble pairs straddle bytes: out byte i = low nibble of b[i], high of b[i+1]
    uint32_t a, e;
    window(kc - 1, &a, &e);
    if (a < e) {
      const uint8_t *sp = key + first / 2 + (a - pos);
      uint32_t c = e - a;
      for (; c > 0; sp += 8) {
        uint64_t x = load8(sp), y = load8(sp + 1);
        uint64_t v = ((x << 4) & 0xF0F0F0F0F0F0F0F0ull) | ((y >> 4) & 0x0F0F0F0F0F0F0F0Full);
        uint32_t t = c < 8 ? c : 8;
        if (t < 8) v &= (1ull << (8 * t)) - 1;
        push(v, t);
        c -= t;
      }
    }
    pos += kc - 1;
  }
  // flush the pending word and zero the block's remaining words
  __device__ __forceinline__ void finish() {
    if (accn) {
      *(uint64_t *)(lane + nw * 512) = acc;
      nw++;
    }
    for (uint32_t w = nw; w < 17; w++) *(uint64_t *)(lane + w * 512) = 0;
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

// Emits the node's full encoding, keeping only the bytes of the current block.
template <typename St>
__device__ __forceinline__ void emit_node(const Pools &P, ref_t r, uint32_t payload, St &st) {
  uint32_t hn;
  uint64_t h = rlp_hdr(payload, 0xC0, 0xF7, &hn);
  st.inline_bytes(h, hn);
  if (ref_type(r) == T_SHORT) {
    const Short &S = P.shrt[ref_idx(r)];
    st.rlp_compact(P.keys + S.key_off, S.key_start, S.key_len - S.term, S.term);
    ChildRef c = child_bytes(P, S.child);
    st.rlp_mem(c.p, c.len);
  } else {
    const Branch &B = P.branch[ref_idx(r)];
    for (int s = 0; s < 17; s++) {
      ref_t c = B.child[s];
      if (c == REF_NULL) {
        st.inline_bytes(0x80, 1);
      } else {
        ChildRef cb = child_bytes(P, c);
        st.rlp_mem(cb.p, cb.len);
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
  for (int s = 0; s < 17; s++) {
    ref_t c = B.child[s];
    pay += c == REF_NULL ? 1u : child_item_len(child_bytes(P, c));
  }
  return pay;
}

__global__ __launch_bounds__(256) void k_hc_hash(Pools P, const ref_t *list, uint32_t n) {
  extern __shared__ __align__(16) uint8_t lds[];
  const uint32_t lane = __lane_id();
  const uint32_t w = threadIdx.x >> 6;
  uint8_t *region = lds + (size_t)w * (17 * 512);
  uint8_t *mine = region + lane * 8;
  uint32_t i = grid_tid();
  bool active = i < n;
  ref_t r = active ? list[i] : REF_NULL;
  uint32_t payload = 0, enc = 0, nblocks = 0;
  if (active) {
    payload = node_payload(P, r);
    enc = rlp_hdr_len(payload) + payload;
    nblocks = enc < 32 ? 1u : enc / 136 + 1;
  }
  uint32_t wave_blocks = nblocks;
  uint32_t perms = enc >= 32 ? nblocks : 0u, bytes = enc;
  for (int off = 32; off > 0; off >>= 1) {
    wave_blocks = max(wave_blocks, (uint32_t)__shfl_xor(wave_blocks, off));
    perms += (uint32_t)__shfl_xor(perms, off);
    bytes += (uint32_t)__shfl_xor(bytes, off);
  }
  if (lane == 0) {
    atomicAdd(&P.cnt->n_perms, (unsigned long long)perms);
    atomicAdd(&P.cnt->enc_bytes, (unsigned long long)bytes);
  }

  uint64_t s[25];
#pragma unroll
  for (int q = 0; q < 25; q++) s[q] = 0;

  for (uint32_t b = 0; b < wave_blocks; b++) {
    bool mine_b = b < nblocks;
    if (mine_b) {
      Blk st(mine, b * 136);
      emit_node(P, r, payload, st);
      st.finish();
      if (enc >= 32 && b + 1 == nblocks) {
        // Keccak padding: 0x01 after the message, 0x80 on the block's last byte
        uint32_t k = enc - st.lo;
        mine[(k >> 3) * 512 + (k & 7)] ^= 0x01;
        mine[16 * 512 + 7] ^= 0x80;
      }
      if (enc >= 32) {
#pragma unroll
        for (int q = 0; q < 17; q++) s[q] ^= *(const uint64_t *)(mine + q * 512);
      }
    }
    if (mine_b && enc >= 32) keccak_f1600(s);
  }
  if (!active) return;
  uint8_t *hash;
  if (ref_type(r) == T_SHORT)
    hash = P.shrt[ref_idx(r)].hash;
  else
    hash = P.branch[ref_idx(r)].hash;
  uint32_t hs;
  if (enc >= 32) {
    uint64_t *h64 = (uint64_t *)hash;
    h64[0] = s[0];
    h64[1] = s[1];
    h64[2] = s[2];
    h64[3] = s[3];
    hs = 32;
  } else {
    for (uint32_t q = 0; q < enc; q++) hash[q] = mine[(q >> 3) * 512 + (q & 7)];
    for (uint32_t q = enc; q < 32; q++) hash[q] = 0;
    hs = enc;
  }
  if (ref_type(r) == T_SHORT) {
    P.shrt[ref_idx(r)].hash_size = (uint8_t)hs;
    P.shrt[ref_idx(r)].dirty = 0;
  } else {
    P.branch[ref_idx(r)].hash_size = (uint8_t)hs;
    P.branch[ref_idx(r)].dirty = 0;
  }
}

__global__ __launch_bounds__(256) void k_hc_roots(Pools P, uint32_t ntries) {
  uint32_t t = grid_tid();
  if (t >= ntries) return;
  TrieHdr &T = P.tries[t];
  ref_t r = T.root;
  uint32_t hs = 0;
  const uint8_t *h = nullptr;
  if (ref_type(r) == T_SHORT) {
    hs = P.shrt[ref_idx(r)].hash_size;
    h = P.shrt[ref_idx(r)].hash;
  } else if (ref_type(r) == T_BRANCH) {
    hs = P.branch[ref_idx(r)].hash_size;
    h = P.branch[ref_idx(r)].hash;
  }
  for (int q = 0; q < 32; q++) T.root_hash[q] = q < (int)hs ? h[q] : 0;
  T.root_hash_size = hs;
}

// Standalone batch Keccak-256: message i is bytes [off[i], off[i] + len[i]) of
// data; digests to out[32 i].  Used to test the permutation and sponge.
__global__ __launch_bounds__(256) void k_keccak_batch(const uint8_t *data, const uint64_t *off,
                                                      const uint32_t *len, uint32_t n, uint8_t *out) {
  extern __shared__ __align__(16) uint8_t lds[];
  const uint32_t lane = __lane_id();
  uint8_t *mine = lds + (size_t)(threadIdx.x >> 6) * (17 * 512) + lane * 8;
  uint32_t i = grid_tid();
  bool active = i < n;
  uint32_t L = active ? len[i] : 0;
  const uint8_t *src = active ? data + off[i] : data;
  uint32_t nblocks = active ? L / 136 + 1 : 0;
  uint32_t wb = nblocks;
  for (int o = 32; o > 0; o >>= 1) wb = max(wb, (uint32_t)__shfl_xor(wb, o));
  uint64_t s[25];
#pragma unroll
  for (int q = 0; q < 25; q++) s[q] = 0;
  for (uint32_t b = 0; b < wb; b++) {
    bool mb = b < nblocks;
    if (mb) {
      Blk st(mine, b * 136);
      st.mem_bytes(src, L);
      st.finish();
      if (b + 1 == nblocks) {
        uint32_t k = L - st.lo;
        mine[(k >> 3) * 512 + (k & 7)] ^= 0x01;
        mine[16 * 512 + 7] ^= 0x80;
      }
#pragma unroll
      for (int q = 0; q < 17; q++) s[q] ^= *(const uint64_t *)(mine + q * 512);
      keccak_f1600(s);
    }
  }
  if (!active) return;
  for (int q = 0; q < 32; q++) out[32 * (size_t)i + q] = (uint8_t)(s[q / 8] >> (8 * (q % 8)));
}

// Node export: encoding length of each node.
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
  Stage<MemSink, 0xffffffffu> st;
  st.sink.out = encs + offs[i];
  st.lo = 0;
  st.pos = 0;
  emit_node(P, r, pay, st);
  const uint8_t *h = ref_type(r) == T_SHORT ? P.shrt[ref_idx(r)].hash : P.branch[ref_idx(r)].hash;
  uint32_t hs = ref_type(r) == T_SHORT ? P.shrt[ref_idx(r)].hash_size : P.branch[ref_idx(r)].hash_size;
  for (int q = 0; q < 32; q++) hashes[32 * (size_t)i + q] = q < (int)hs ? h[q] : 0;
}

}  // namespace

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

size_t hc_lds_bytes() { return 4 * 17 * 512; }

hipError_t hc_classify(const Pools &P, const ref_t *list, uint32_t n, uint32_t *hist, uint32_t *cursor,
                       ref_t *out, void *temp, hipStream_t s) {
  hipError_t e = hipMemsetAsync(hist, 0, sizeof(uint32_t) * kHcBins, s);
  if (e != hipSuccess) return e;
  if (n == 0) return hipSuccess;
  unsigned tiles = (unsigned)((n + kHcTile - 1) / kHcTile);
  k_hc_hist<<<tiles, kB, 0, s>>>(P, list, n, hist);
  GMPT_LAUNCH_CHECK();
  e = scan_exclusive_u32(hist, cursor, kHcBins, nullptr, temp, s);
  if (e != hipSuccess) return e;
  k_hc_scatter<<<tiles, kB, 0, s>>>(P, list, n, cursor, out);
  return hipGetLastError();
}

hipError_t hc_hash_range(const Pools &P, const ref_t *list, uint32_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_hc_hash<<<nblk(n), kB, hc_lds_bytes(), s>>>(P, list, n);
  return hipGetLastError();
}

hipError_t hc_roots(const Pools &P, uint32_t ntries, hipStream_t s) {
  k_hc_roots<<<nblk(ntries), kB, 0, s>>>(P, ntries);
  return hipGetLastError();
}

hipError_t keccak_batch(const uint8_t *data, const uint64_t *off, const uint32_t *len, uint32_t n, uint8_t *out,
                        hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_keccak_batch<<<nblk(n), kB, hc_lds_bytes(), s>>>(data, off, len, n, out);
  return hipGetLastError();
}

}  // namespace gmpt
