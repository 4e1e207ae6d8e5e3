// phasenu.hip -- PhaseNU: sort-based batch insert into empty or existing tries.
//
// Reference: puts_2phase_with_valuehp (mpt-with-compress/include/mpt/gpu_mpt.cuh:
// 1262-1399) runs split (gpu_mpt_kernels.cuh:3592-3673), a per-nibble full
// expansion insert (:3823-4056) and compress (:4420-4666).  This version sorts
// the batch's nibble paths instead, so each trie node is built once:
//   1. pack hex keys into the key arena; sort (trie, 16-nibble window) pairs
//   2. adjacent LCPs, dedup keeping the last occurrence (last writer wins,
//      matching the CPU MPT's sequential puts_baseline, cpu_mpt.cuh:359-378)
//   3. descend each unique key into the existing trie to its stop point
//   4. per stop point ("group") build the new subtree from the LCP array: each
//      distinct LCP interval is a branch, gaps between depths are extension
//      short nodes, each key a leaf short node; a short node that new keys
//      split joins its group as one more element ("virtual" element)
// The result is the same trie the sequential insert produces.
#include <limits.h>
#include "hc_leaf.h"
#include "phasenu.h"
#include "prims.h"

namespace gmpt {

namespace {

constexpr int kB = 256;
constexpr uint32_t kDirtyStage = 4096;  // staged dirty-list entries per block
static_assert(SC_NE == SC_NE_INDEX, "the replay check reads the element count at SC_NE");

__device__ __forceinline__ uint32_t grid_tid() { return blockIdx.x * blockDim.x + threadIdx.x; }

inline unsigned nblk(size_t n) { return (unsigned)((n + kB - 1) / kB); }

// A key stored in the key arena: nibble j at absolute path position j, the
// terminator at position len-1.
struct KeyRef {
  const uint8_t *p;
  uint32_t len;
  __device__ __forceinline__ uint32_t nib(uint32_t j) const { return j + 1 == len ? kNibbleT : packed_nibble(p, j); }
};

constexpr uint64_t kKrefOff = (1ull << 48) - 1;
__device__ __forceinline__ KeyRef key_of(const NuArgs &a, uint32_t idx) {
  const uint64_t r = a.kref[idx];
  KeyRef k;
  k.p = a.P.keys + a.kbase + (r & kKrefOff);
  k.len = (uint32_t)(r >> 48);
  return k;
}

// Nibble at absolute position j (inside the node's key range) of short node S.
__device__ __forceinline__ uint32_t short_nib(const Pools &P, const Short &S, uint32_t j) {
  uint32_t rel = j - S.depth;
  if (S.term && rel + 1 == S.key_len) return kNibbleT;
  return packed_nibble(P.keys + S.key_off, S.key_start + rel);
}

// ------------------------------------------------------------- 1. keys ----

__device__ __forceinline__ bool key_range_ok(const NuArgs &a, uint32_t i) {
  int32_t s = a.kidx[2 * i], e = a.kidx[2 * i + 1];
  return s >= 0 && e >= s && (uint64_t)e < a.kh_bytes && e - s + 1 <= kMaxKeyNibbles;
}

// the first c (0..8) bytes at p (upper bytes zero)
__device__ __forceinline__ uint64_t ldn(const uint8_t *p, uint32_t c) {
  if (c == 0) return 0;
  uintptr_t a = (uintptr_t)p;
  uint32_t sh = (uint32_t)(a & 7);
  const uint64_t *q = (const uint64_t *)(a - sh);
  uint64_t v = q[0] >> (8 * sh);
  if (sh + c > 8) v |= q[1] << (64 - 8 * sh);
  return c >= 8 ? v : v & ((1ull << (8 * c)) - 1);
}
// 8 hex nibble bytes -> 4 packed bytes (high nibble first), memory order
__device__ __forceinline__ uint32_t pack8(uint64_t h) {
  uint64_t t = ((h & 0x0F000F000F000F00ull) >> 8) | ((h & 0x000F000F000F000Full) << 4);
  return (uint32_t)((t & 0xFFull) | ((t >> 8) & 0xFF00ull) | ((t >> 16) & 0xFF0000ull) | ((t >> 24) & 0xFF000000ull));
}

// Hex-key staging: the block's keys are read from LDS when their bytes span
// at most kKeyStage bytes (keys stored in batch order), so the block's HBM
// reads are coalesced 16-byte loads instead of one strided stream per lane.
constexpr uint32_t kKeyStage = 17408;

struct KeyStager {
  uintptr_t base = 0;  // device address of lds[0] (16-byte aligned), 0 when not staged
  const uint8_t *lds = nullptr;
  // pointer to hex byte `off` of the batch's key buffer
  __device__ __forceinline__ const uint8_t *at(const NuArgs &a, int64_t off) const {
    return base ? lds + ((uintptr_t)(a.kh + off) - base) : a.kh + off;
  }
};

// Every thread of the block must call this (it synchronises the block).
__device__ __forceinline__ KeyStager stage_keys(const NuArgs &a, uint8_t *lds) {
  __shared__ uint64_t red_lo[4], red_hi[4];
  uint32_t i = grid_tid();
  uint64_t lo = ~0ull, hi = 0;
  if (i < a.n && key_range_ok(a, i)) {
    lo = (uint64_t)a.kidx[2 * i];
    hi = (uint64_t)a.kidx[2 * i + 1];
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo = min(lo, (uint64_t)__shfl_xor((unsigned long long)lo, off));
    hi = max(hi, (uint64_t)__shfl_xor((unsigned long long)hi, off));
  }
  if (__lane_id() == 0) {
    red_lo[threadIdx.x >> 6] = lo;
    red_hi[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  lo = min(min(red_lo[0], red_lo[1]), min(red_lo[2], red_lo[3]));
  hi = max(max(red_hi[0], red_hi[1]), max(red_hi[2], red_hi[3]));
  KeyStager st;
  st.lds = lds;
  if (lo == ~0ull) return st;
  // aligned 16-byte words covering [kh + lo, kh + hi]: they stay inside the
  // memory pages holding those bytes
  uintptr_t b = ((uintptr_t)(a.kh + lo)) & ~(uintptr_t)15;
  uintptr_t e = ((uintptr_t)(a.kh + hi) + 16) & ~(uintptr_t)15;
  if (e - b > kKeyStage) return st;
  uint32_t nw = (uint32_t)((e - b) / 16);
  // every load of the thread issued before its LDS stores (a store waits for
  // its load, so a rolled loop would make one memory round trip per word)
  constexpr uint32_t kStagePer = (kKeyStage / 16 + 255) / 256;
  uint32_t v[4 * kStagePer];
#pragma unroll
  for (uint32_t q = 0; q < kStagePer; q++) {
    const uint32_t k = threadIdx.x + 256 * q;
    uint4 x = make_uint4(0, 0, 0, 0);
    if (k < nw) x = ((const uint4 *)b)[k];
    v[4 * q] = x.x;
    v[4 * q + 1] = x.y;
    v[4 * q + 2] = x.z;
    v[4 * q + 3] = x.w;
  }
#pragma unroll
  for (uint32_t q = 0; q < kStagePer; q++) {
    const uint32_t k = threadIdx.x + 256 * q;
    if (k < nw) ((uint4 *)lds)[k] = make_uint4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
  }
  __syncthreads();
  st.base = b;
  return st;
}

// Packed slot bytes of a key of L hex bytes (terminator included): two
// nibbles per byte, padded to whole 8-byte words, at least one zero byte.
__device__ __forceinline__ uint32_t slot_bytes(uint32_t L) { return (L / 2 + 8) & ~7u; }

// Writes the key's nibbles (nn = L - 1 of them, the terminator left out) two
// per byte, high nibble first, zero past nn, over slot_bytes(L) bytes.
__device__ __forceinline__ void pack_key(uint64_t *dst, const uint8_t *k, uint32_t L) {
  const uint32_t nb = slot_bytes(L);
  const uint32_t nn = L - 1;
  for (uint32_t b = 0; b < nb; b += 8) {
    uint32_t o = 2 * b;  // first nibble of this 8-byte word
    uint64_t h0 = o < nn ? ldn(k + o, min(8u, nn - o)) : 0ull;
    uint64_t h1 = o + 8 < nn ? ldn(k + o + 8, min(8u, nn - o - 8)) : 0ull;
    dst[b / 8] = (uint64_t)pack8(h0) | ((uint64_t)pack8(h1) << 32);
  }
}

// 8 hex nibble bytes (0x0N each, memory order) -> 4 packed bytes, high nibble
// first: per 32-bit half, byte pairs merge by shifts, then one byte permute
// picks bytes 0 and 2 of each half.
__device__ __forceinline__ uint32_t pack8_perm(uint64_t h) {
  const uint32_t lo = (uint32_t)h, hi = (uint32_t)(h >> 32);
  const uint32_t ylo = (lo << 4) | (lo >> 8), yhi = (hi << 4) | (hi >> 8);
  return __builtin_amdgcn_perm(yhi, ylo, 0x06040200u);
}

// A key of at most kFastKeyBytes hex bytes held in registers: the aligned
// words holding its bytes (no other word is read), and chunk j = its bytes
// [8j, 8j + 8) (zero past the key).
constexpr uint32_t kFastKeyBytes = 72;
constexpr int kFastKeyWords = (int)((7 + kFastKeyBytes + 7) / 8);  // aligned words a key can touch
constexpr int kFastKeyChunks = (int)(kFastKeyBytes / 8);
struct KeyChunks {
  uint64_t c[kFastKeyChunks + 1];
  __device__ __forceinline__ void load(const uint8_t *k, uint32_t L) {
    const uintptr_t a = (uintptr_t)k;
    const uint32_t sh = (uint32_t)(a & 7);
    const uint64_t *q = (const uint64_t *)(a - sh);
    const uint32_t nw = (sh + L + 7) / 8;
    uint64_t w[kFastKeyWords + 1];
#pragma unroll
    for (int j = 0; j <= kFastKeyWords; j++) w[j] = (uint32_t)j < nw ? q[j] : 0ull;
#pragma unroll
    for (int j = 0; j <= kFastKeyChunks; j++) {
      const uint64_t x = sh ? (w[j] >> (8 * sh)) | (w[j + 1] << (64 - 8 * sh)) : w[j];
      const int32_t left = (int32_t)L - 8 * j;  // key bytes from this chunk's first
      c[j] = left >= 8 ? x : left > 0 ? x & ((1ull << (8 * left)) - 1) : 0ull;
    }
  }
  // byte b (b < kFastKeyBytes)
  __device__ __forceinline__ uint32_t byte(uint32_t b) const {
    uint64_t v = 0;
#pragma unroll
    for (int j = 0; j <= kFastKeyChunks; j++) v = (uint32_t)j == b / 8 ? c[j] : v;
    return (uint32_t)(v >> (8 * (b & 7))) & 0xffu;
  }
};

// Validates every key, and reduces per block the common prefix with key 0
// and the longest key (no global atomics: one slot per block).  A key as long
// as key 0 is also packed at once to slot i of a layout with equal slots: when
// every key has that length (bit 31 of blk_len stays clear in every block),
// the packing pass only builds the sort windows.
__global__ __launch_bounds__(256) void k_keys_scan(NuArgs a, uint32_t *blk_w, uint32_t *blk_len) {
  __shared__ uint32_t sw[4], sl[4];
  __shared__ __align__(16) uint8_t lds[kKeyStage + 16];
  const KeyStager stg = stage_keys(a, lds);
  uint32_t i = grid_tid();
  uint32_t lcp0 = UINT_MAX, len = 0, odd = 0;
  uint64_t sbytes = 0;  // the key's packed slot size
  const uint32_t L0 = key_range_ok(a, 0) ? (uint32_t)(a.kidx[1] - a.kidx[0] + 1) : 0u;
  if (i < a.n) {
    bool ok = key_range_ok(a, i);
    int32_t L = 1;
    bool packed = false;  // a short key checked and packed from registers
    if (ok) {
      int32_t s = a.kidx[2 * i], e = a.kidx[2 * i + 1];
      L = e - s + 1;
      const uint8_t *k = stg.at(a, s);
      uint32_t nn = (uint32_t)L - 1;  // nibbles before the terminator
      if ((uint32_t)L <= kFastKeyBytes) {
        KeyChunks kc;
        kc.load(k, (uint32_t)L);
        uint64_t hi = 0;  // high nibbles of the bytes before the terminator
#pragma unroll
        for (int j = 0; j < kFastKeyChunks; j++) {
          const int32_t left = (int32_t)nn - 8 * j;
          hi |= left >= 8 ? kc.c[j] : left > 0 ? kc.c[j] & ((1ull << (8 * left)) - 1) : 0ull;
        }
        ok = kc.byte(nn) == kNibbleT && (hi & 0xF0F0F0F0F0F0F0F0ull) == 0;
        // the same slot layout as pack_key: nibble pairs, zero past nn
        if (ok && (uint32_t)L == L0 && a.kbase + (uint64_t)(i + 1) * slot_bytes((uint32_t)L) <= a.key_cap) {
          uint64_t *dst = (uint64_t *)(a.P.keys + a.kbase + (uint64_t)i * slot_bytes((uint32_t)L));
          const uint32_t nbw = slot_bytes((uint32_t)L) / 8;
          uint64_t first = 0;  // the slot's first word: nibbles 0 .. 15
#pragma unroll
          for (int b = 0; b < (kFastKeyChunks + 1) / 2; b++) {
            if ((uint32_t)b >= nbw) break;
            uint64_t h0 = kc.c[2 * b], h1 = kc.c[2 * b + 1];
            const int32_t l0 = (int32_t)nn - 16 * b, l1 = l0 - 8;  // real nibbles left in each chunk
            h0 = l0 >= 8 ? h0 : l0 > 0 ? h0 & ((1ull << (8 * l0)) - 1) : 0ull;
            h1 = l1 >= 8 ? h1 : l1 > 0 ? h1 & ((1ull << (8 * l1)) - 1) : 0ull;
            dst[b] = (uint64_t)pack8_perm(h0) | ((uint64_t)pack8_perm(h1) << 32);
            if (b == 0) first = dst[b];
          }
          packed = true;
          // the sort window of a batch with no common prefix (W = 0), as
          // k_keys_pack builds it from the slot: that pass skips such a batch
          if (a.sk0) {
            a.sk0[i] = __builtin_bswap64(first);
            a.sv0[i] = i;
          }
        }
      } else {
        ok = k[nn] == kNibbleT;
        for (uint32_t o = 0; ok && o < nn; o += 8) {
          uint64_t w = ldn(k + o, min(8u, nn - o));
          ok = (w & 0xF0F0F0F0F0F0F0F0ull) == 0;
        }
      }
      // common prefix with key 0 (the terminator is a symbol of its own)
      if (ok && key_range_ok(a, 0)) {
        int32_t s0 = a.kidx[0], e0 = a.kidx[1];
        const uint8_t *k0 = a.kh + s0;
        uint32_t mn = min((uint32_t)L, (uint32_t)(e0 - s0 + 1));
        uint32_t p = 0;
        for (; p < mn; p += 8) {
          uint32_t c = min(8u, mn - p);
          uint64_t x = ldn(k + p, c) ^ ldn(k0 + p, c);
          if (x) {
            p += (uint32_t)(__ffsll((unsigned long long)x) - 1) / 8;
            break;
          }
        }
        lcp0 = min(p, mn);
      } else {
        lcp0 = 0;
      }
    }
    if (ok && a.tid && (uint32_t)a.tid[i] >= a.ntries) ok = false;
    if (ok) {
      int64_t vs = a.vidx[2 * i], ve = a.vidx[2 * i + 1];
      ok = vs >= 0 && ve >= vs - 1 && (uint64_t)(ve + 1) <= a.v_bytes && ve - vs + 1 < (1ll << 31);
    }
    if (!ok) {
      raise_error(a.P.cnt, ERR_BAD_KEY, i);
      L = 1;
      lcp0 = 0;
    }
    len = (uint32_t)L;
    a.klen[i] = len;
    const uint32_t slot = slot_bytes(len);
    sbytes = slot;
    // a rejected key turns the equal-slot layout off: the general packing
    // writes its kref (the equal-slot pass only writes those of keys it packs)
    odd = !ok || len != L0;
    if (ok && !odd && a.kbase + (uint64_t)(i + 1) * slot <= a.key_cap) {
      const uint64_t koff = (uint64_t)i * slot;
      if (!packed) pack_key((uint64_t *)(a.P.keys + a.kbase + koff), stg.at(a, a.kidx[2 * i]), len);
      a.kref[i] = koff | ((uint64_t)len << 48);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    lcp0 = min(lcp0, (uint32_t)__shfl_xor(lcp0, off));
    len = max(len, (uint32_t)__shfl_xor(len, off));
    odd |= (uint32_t)__shfl_xor(odd, off);
    sbytes += (uint64_t)__shfl_xor((unsigned long long)sbytes, off);
  }
  __shared__ uint64_t sb[4];
  if (__lane_id() == 0) {
    sw[threadIdx.x >> 6] = lcp0;
    sl[threadIdx.x >> 6] = len | (odd << 31);
    sb[threadIdx.x >> 6] = sbytes;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // the block's slot bytes: k_keys_reduce turns these into block offsets
    a.kpoff[blockIdx.x] = sb[0] + sb[1] + sb[2] + sb[3];
    blk_w[blockIdx.x] = min(min(sw[0], sw[1]), min(sw[2], sw[3]));
    const uint32_t m = sl[0] | sl[1] | sl[2] | sl[3];
    const uint32_t l = max(max(sl[0] & 0xFFFFu, sl[1] & 0xFFFFu), max(sl[2] & 0xFFFFu, sl[3] & 0xFFFFu));
    blk_len[blockIdx.x] = l | (m & 0x80000000u);
  }
}

__global__ __launch_bounds__(256) void k_keys_reduce(NuArgs a, const uint32_t *blk_w, const uint32_t *blk_len,
                                                     uint32_t nb) {
  __shared__ uint32_t sw[4], sl[4], so[4];
  __shared__ uint64_t s64[4];
  uint32_t w = UINT_MAX, l = 0, odd = 0;
  for (uint32_t b0 = threadIdx.x; b0 < nb; b0 += 8 * blockDim.x) {
#pragma unroll
    for (uint32_t q = 0; q < 8; q++) {  // eight blocks' loads in flight
      const uint32_t b = b0 + q * blockDim.x;
      if (b < nb) {
        const uint32_t bw = blk_w[b], bl = blk_len[b];
        w = min(w, bw);
        l = max(l, bl & 0xFFFFu);
        odd |= bl >> 31;
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    w = min(w, (uint32_t)__shfl_xor(w, off));
    l = max(l, (uint32_t)__shfl_xor(l, off));
    odd |= (uint32_t)__shfl_xor(odd, off);
  }
  if (__lane_id() == 0) {
    sw[threadIdx.x >> 6] = w;
    sl[threadIdx.x >> 6] = l;
    so[threadIdx.x >> 6] = odd;
  }
  __syncthreads();
  // every key as long as key 0, and every one of them packed: the offsets
  // are i * slot and the packing pass only builds the sort windows
  const uint32_t L0 = key_range_ok(a, 0) ? (uint32_t)(a.kidx[1] - a.kidx[0] + 1) : 0u;
  const uint64_t ktot = (uint64_t)a.n * slot_bytes(L0);
  const bool uni = (so[0] | so[1] | so[2] | so[3]) == 0 && a.n > 0 && L0 > 0 && a.kbase + ktot <= a.key_cap;
  uint64_t kbytes = ktot;
  if (!uni) {  // uniform over the block
    // exclusive scan of the blocks' slot bytes in place (each thread a
    // contiguous run of blocks): the packing pass adds a block-local scan
    const uint32_t per = (nb + blockDim.x - 1) / blockDim.x;
    const uint32_t b0 = min(nb, threadIdx.x * per), b1 = min(nb, b0 + per);
    uint64_t part = 0;
    for (uint32_t b = b0; b < b1; b++) part += a.kpoff[b];
    uint64_t pre = block_exclusive_scan_256<uint64_t>(part, s64, &kbytes);
    for (uint32_t b = b0; b < b1; b++) {
      const uint64_t v = a.kpoff[b];
      a.kpoff[b] = pre;
      pre += v;
    }
  }
  if (threadIdx.x == 0) {
    a.scal[SC_W] = min(min(sw[0], sw[1]), min(sw[2], sw[3]));
    a.scal[SC_MAXLEN] = max(max(sl[0], sl[1]), max(sl[2], sl[3]));
    a.scal[SC_UNIFORM] = uni;
    *(uint64_t *)(a.scal + SC_KTOT) = kbytes;
  }
}

// Packs nibbles two per byte into the key arena and builds the sort window:
// 16 nibbles from the batch's common prefix W, terminator and padding as 0.
__global__ __launch_bounds__(256) void k_keys_pack(NuArgs a) {
  __shared__ __align__(16) uint8_t lds[kKeyStage + 16];
  if (a.scal[SC_UNIFORM]) {
    // keys already packed in equal slots: the window is read from the packed
    // bytes (nibble W sits in 8-byte word W / 16 of the slot, and the slot
    // holds byte W / 2 since W <= L)
    const uint32_t i = grid_tid();
    const uint32_t W = a.scal[SC_W];
    // every key took k_keys_scan's register path (equal lengths of at most
    // kFastKeyBytes), which wrote the windows for W = 0
    if (W == 0 && a.scal[SC_MAXLEN] <= kFastKeyBytes) return;
    if (i >= a.n || !a.sk0) return;
    const uint32_t L = a.klen[i];
    const uint32_t slot = slot_bytes(L);
    const uint64_t *src = (const uint64_t *)(a.P.keys + a.kbase + (uint64_t)i * slot);
    const uint32_t q = W / 16, d = W % 16;
    const uint64_t hi = __builtin_bswap64(src[q]);
    const uint64_t lo = (q + 1) * 8 < slot ? __builtin_bswap64(src[q + 1]) : 0ull;
    a.sk0[i] = d ? (hi << (4 * d)) | (lo >> (64 - 4 * d)) : hi;
    a.sv0[i] = i;
    return;
  }
  const KeyStager stg = stage_keys(a, lds);
  uint32_t i = grid_tid();
  // the key's slot: its block's offset plus the slots before it in the block
  __shared__ uint64_t s64[4];
  const uint32_t Li = i < a.n ? a.klen[i] : 0u;
  uint64_t btot;
  const uint64_t koff =
      a.kpoff[blockIdx.x] + block_exclusive_scan_256<uint64_t>(i < a.n ? slot_bytes(Li) : 0u, s64, &btot);
  if (i >= a.n) return;
  // The packed keys must fit the arena.  The host reserves it from the key
  // buffer's size, which index pairs that overlap or repeat a key range can
  // exceed: then no key is packed, every key reads as the empty key (its
  // terminator only, so every later read stays at the batch's first arena
  // bytes) and the call fails with an arena error; the host grows the arena
  // to the total and runs the insert again.
  if (a.kbase + *(const uint64_t *)(a.scal + SC_KTOT) > a.key_cap) {
    a.kref[i] = 1ull << 48;
    if (a.sk0) {
      a.sk0[i] = 0;
      a.sv0[i] = i;
    }
    if (i == 0) raise_error(a.P.cnt, ERR_POOL, kErrArgKeyArena);
    return;
  }
  const uint32_t L = Li;
  const uint8_t *k = key_range_ok(a, i) ? stg.at(a, a.kidx[2 * i]) : a.kh;
  a.kref[i] = koff | ((uint64_t)L << 48);
  pack_key((uint64_t *)(a.P.keys + a.kbase + koff), k, L);
  const uint32_t nn = L - 1;  // real nibbles (0 for a rejected key)
  if (!a.sk0) return;  // LockNU packs keys without sorting them
  uint32_t W = a.scal[SC_W];
  uint64_t h0 = W < nn ? ldn(k + W, min(8u, nn - W)) : 0ull;
  uint64_t h1 = W + 8 < nn ? ldn(k + W + 8, min(8u, nn - W - 8)) : 0ull;
  a.sk0[i] = ((uint64_t)__builtin_bswap32(pack8(h0)) << 32) | __builtin_bswap32(pack8(h1));
  a.sv0[i] = i;
}

// Full-key sort support: word `w` (16 nibbles from W + 16w) of key sv[j], or
// the key length for w < 0.
__global__ __launch_bounds__(256) void k_gather_word(NuArgs a, const uint32_t *sv, uint64_t *sk, int w) {
  uint32_t j = grid_tid();
  if (j >= a.n) return;
  uint32_t i = sv[j];
  uint32_t L = a.klen[i];
  if (w < 0) {
    sk[j] = L;
    return;
  }
  const uint8_t *k = a.kh + (key_range_ok(a, i) ? a.kidx[2 * i] : 0);
  uint32_t nn = L - 1;
  uint32_t W = a.scal[SC_W] + 16u * (uint32_t)w;
  uint64_t win = 0;
  for (int q = 0; q < 16; q++) {
    uint32_t jj = W + q;
    uint64_t v = jj < nn ? (uint64_t)k[jj] : 0ull;
    win |= v << (60 - 4 * q);
  }
  sk[j] = win;
}

__global__ __launch_bounds__(256) void k_gather_trie(NuArgs a, const uint32_t *sv, uint32_t *tk) {
  uint32_t j = grid_tid();
  if (j >= a.n) return;
  tk[j] = a.tid ? (uint32_t)a.tid[sv[j]] : 0u;
}

// ----------------------------------------------------- 2. adjacent/dedup --

__device__ __forceinline__ uint32_t trie_of(const NuArgs &a, uint32_t idx) {
  return a.tid ? (uint32_t)a.tid[idx] : 0u;
}

// Bit gather / scatter over a mask that is the same for every lane (the loop
// runs once per mask bit, lowest first).
__device__ __forceinline__ uint64_t gather_bits(uint64_t x, uint64_t m) {
  uint64_t r = 0;
  for (uint32_t k = 0; m; k++) {
    const uint64_t b = m & (0ull - m);
    r |= (uint64_t)((x & b) != 0) << k;
    m ^= b;
  }
  return r;
}
__device__ __forceinline__ uint64_t scatter_bits(uint64_t x, uint64_t m) {
  uint64_t r = 0;
  while (m) {
    const uint64_t b = m & (0ull - m);
    if (x & 1) r |= b;
    x >>= 1;
    m ^= b;
  }
  return r;
}

// Offset of this block's first entry from per-block counts cnt (one per
// 256-entry block of the launch): with pre, cnt already holds the exclusive
// scan of the counts; otherwise the block adds up the counts of the blocks
// before it (few blocks: kSelfSumBlocks at most).  Every thread of the block
// calls it (it synchronises the block unless pre).
constexpr uint32_t kSelfSumBlocks = 4096;
__device__ __forceinline__ uint32_t block_offset(const uint32_t *cnt, bool pre, uint32_t *sh) {
  if (pre) return cnt[blockIdx.x];
  uint32_t part = 0;  // all of the thread's loads in flight at once
#pragma unroll
  for (uint32_t q = 0; q < kSelfSumBlocks / 256; q++) {
    const uint32_t t = threadIdx.x + 256 * q;
    part += t < blockIdx.x ? cnt[t] : 0u;
  }
  uint32_t tot;
  block_exclusive_scan_256<uint32_t>(part, sh, &tot);
  return tot;
}

// Adjacent keys of the sorted order: LCP, duplicate and new-trie flags.
// With cmask != 0 the keys are windows compressed to the bits of cmask
// (k_compress_windows): the XOR of two is scattered back to window bits
// before its leading zeros give the first differing nibble.
__global__ __launch_bounds__(256) void k_adjacent(NuArgs a, const uint64_t *sk, const uint32_t *sv, int full,
                                                  uint64_t cmask) {
  const uint32_t j = grid_tid();
  const bool active = j < a.n;
  // with every key as long as key 0, packed in equal slots, a key's place is
  // its input index times the slot size: no load
  const uint32_t Lu = a.scal[SC_UNIFORM] ? a.scal[SC_MAXLEN] : 0u;
  const uint64_t su = slot_bytes(Lu);
  const auto kref_of = [&](uint32_t i) -> uint64_t {
    return Lu ? (uint64_t)i * su | (uint64_t)Lu << 48 : a.kref[i];
  };
  // key j's place and trie; key j - 1's come from the lane before (one
  // random load per key instead of two), or a load on a wave's first lane
  const uint32_t ib = active ? sv[j] : 0u;
  const uint64_t rb = active ? kref_of(ib) : 0ull;
  const uint32_t tb = active ? trie_of(a, ib) : 0u;
  uint64_t ra = (uint64_t)__shfl_up((long long)rb, 1);
  uint32_t ta = (uint32_t)__shfl_up((int)tb, 1);
  uint8_t fl = 0, l = 0;
  if (active) {
  if (a.skref) a.skref[j] = rb;
  if (__lane_id() == 0 && j > 0) {
    const uint32_t ia = sv[j - 1];
    ra = kref_of(ia);
    ta = trie_of(a, ia);
  }
  if (j == 0) {
    fl = 2;
  } else {
    if (ta != tb) {
      fl = 2;
    } else {
      KeyRef ka, kb;
      ka.p = a.P.keys + a.kbase + (ra & kKrefOff);
      ka.len = (uint32_t)(ra >> 48);
      kb.p = a.P.keys + a.kbase + (rb & kKrefOff);
      kb.len = (uint32_t)(rb >> 48);
      uint32_t W = a.scal[SC_W];
      uint32_t mn = min(ka.len, kb.len);
      uint32_t p;
      bool ident;
      uint64_t x = full ? 0ull : (sk[j - 1] ^ sk[j]);
      if (x != 0 && cmask != 0) x = scatter_bits(x >> (64 - __popcll(cmask)), cmask);
      if (x != 0) {
        p = W + (uint32_t)(__clzll((long long)x) >> 2);
        if (ka.len != kb.len) p = min(p, mn - 1);
        ident = false;
      } else {
        // windows agree (or full mode): compare nibble by nibble; with
        // different lengths the shorter key's terminator (0 in the window)
        // differs from the longer key's nibble at mn - 1
        p = full ? 0u : min(W + 16u, mn);
        if (ka.len != kb.len) p = min(p, mn - 1);
        while (p < mn && ka.nib(p) == kb.nib(p)) p++;
        ident = (p == mn) && ka.len == kb.len;
        if (!ident && !full) atomicOr(&a.P.cnt->flag_fallback, 1u);
      }
      fl = ident ? 1 : 0;
      l = (uint8_t)min(p, 255u);
    }
  }
  a.lcp[j] = l;
  a.flags[j] = fl;
  // the unique-order entries for a batch with no equal keys, where key j is
  // unique key j (k_unique rewrites them, compacted, when SC_DUP is set)
  a.u_lcp[j] = l;
  a.u_first[j] = (fl >> 1) & 1;
  a.u_trie[j] = tb;
  a.u_idx[j] = ib;
  if (a.skref) {
    a.u_kref[j] = rb;
    if (a.u_win_ok) a.u_win[j] = sk[j];
  }
  if (fl & 1) atomicOr(&a.scal[SC_DUP], 1u);
  }
  // the block's run starts (keys that are no duplicate of the one before):
  // k_unique numbers the runs from these counts
  const uint32_t starts = (uint32_t)__syncthreads_count(active && (fl & 1) == 0);
  if (threadIdx.x == 0) a.run[blockIdx.x] = starts;
}

// After the inclusive scan of run starts: run r spans sorted positions with
// run[j] == r+1.  The first position carries the LCP to the previous distinct
// key, the last position the surviving (last-written) input index.
__global__ __launch_bounds__(256) void k_unique(NuArgs a, const uint64_t *sk, const uint32_t *sv,
                                                const uint32_t *starts, int pre) {
  // no equal keys: k_adjacent wrote the unique-order entries in place
  if (a.scal[SC_DUP] == 0) {  // uniform over the grid
    if (grid_tid() == 0) a.scal[SC_M_UNIQ] = a.n;
    return;
  }
  __shared__ uint32_t sh[4];
  uint32_t j = grid_tid();
  const bool first = j < a.n && (a.flags[j] & 1) == 0;
  // the run starts before this one (block offset from the per-block counts
  // of k_adjacent, then a block-local scan): a run start's count is its run
  // index; later positions of the run count that start already
  uint32_t btot;
  const uint32_t excl = block_offset(starts, pre != 0, sh) + block_exclusive_scan_256<uint32_t>(first, sh, &btot);
  if (j >= a.n) return;
  bool last = (j + 1 == a.n) || ((a.flags[j + 1] & 1) == 0);
  uint32_t r = first ? excl : excl - 1u;
  if (first) {
    a.u_lcp[r] = a.lcp[j];
    a.u_first[r] = (a.flags[j] >> 1) & 1;
    a.u_trie[r] = trie_of(a, sv[j]);
  }
  if (last) {
    a.u_idx[r] = sv[j];
    if (a.skref) {
      a.u_kref[r] = a.skref[j];
      if (a.u_win_ok) a.u_win[r] = sk[j];
    }
  }
  if (j + 1 == a.n) a.scal[SC_M_UNIQ] = r + 1;
}

// --------------------------------------------------- 3. descend + groups --

__global__ __launch_bounds__(256) void k_descend(NuArgs a) {
  __shared__ ref_t dbuf[kDirtyStage];
  __shared__ uint32_t dcnt;
  if (threadIdx.x == 0) dcnt = 0;
  __syncthreads();
  DirtyStage ds{dbuf, &dcnt, kDirtyStage};
  uint32_t i = grid_tid();
  uint32_t m = a.scal[SC_M_UNIQ];
  bool active = i < m;
  const Pools &P = a.P;
  Stop st;
  st.stop = REF_NULL;
  st.attach = REF_NULL;
  st.base = 0;
  st.div = 0;
  st.kind = STOP_ROOT_EMPTY;
  st.slot = 0;
  st.less = 0;
  st.pad = 0;
  KeyRef k;
  ref_t node = REF_NULL, parent = REF_NULL;
  uint32_t pos = 0, pslot = 0;
  bool done = !active;
  k.p = nullptr;
  k.len = 0;
  if (active) {
    uint32_t idx = a.u_idx[i];
    uint32_t t = a.u_trie[i];
    if (t < a.ntries) {
      node = P.tries[t].root;
      if (node != REF_NULL) k = key_of(a, idx);  // an empty trie needs no key bytes
    } else {
      done = true;
      raise_error(P.cnt, ERR_BAD_KEY, idx);
    }
  }
  for (int guard = 0; guard < 2 * kMaxKeyNibbles + 4 && !done; guard++) {
    bool want_mark = false;
    ref_t mark_ref = REF_NULL;
    if (!done) {
      uint32_t ty = ref_type(node);
      if (node == REF_NULL) {
        if (parent == REF_NULL) {
          st.kind = STOP_ROOT_EMPTY;
          st.base = 0;
        } else {
          st.kind = STOP_BRANCH_SLOT;
          st.attach = parent;
          st.slot = (uint8_t)pslot;
          st.base = (uint16_t)pos;
        }
        done = true;
      } else if (ty == T_BRANCH) {
        const Branch &B = P.branch[ref_idx(node)];
        want_mark = true;
        mark_ref = node;
        if (pos >= k.len || B.depth != pos) {
          raise_error(P.cnt, ERR_INTERNAL, 12);
          done = true;
          break;
        }
        uint32_t nib = k.nib(pos);
        parent = node;
        pslot = nib;
        node = B.child[nib];
        pos += 1;
      } else if (ty == T_SHORT) {
        const ref_t cur = node;
        const Short &S = P.shrt[ref_idx(cur)];
        uint32_t ml = S.key_len;
        uint32_t p = 0;
        while (p < ml && pos + p < k.len && short_nib(P, S, pos + p) == k.nib(pos + p)) p++;
        if (p == ml) {
          want_mark = true;
          mark_ref = cur;
          if (ref_type(S.child) == T_VALUE) {
            st.kind = STOP_LEAF_UPDATE;
            st.stop = cur;
            st.attach = parent;
            st.slot = (uint8_t)pslot;
            st.base = (uint16_t)pos;
            done = true;
          } else {
            parent = cur;
            pslot = 0;
            node = S.child;
            pos += ml;
          }
        } else if (pos + p >= k.len) {
          raise_error(P.cnt, ERR_INTERNAL, 11);
          done = true;
        } else {
          uint32_t d = pos + p;
          uint32_t kn = k.nib(d), sn = short_nib(P, S, d);
          st.kind = STOP_SHORT_SPLIT;
          st.stop = cur;
          st.attach = parent;
          st.slot = (uint8_t)pslot;
          st.base = (uint16_t)pos;
          st.div = (uint16_t)d;
          st.less = (kn == kNibbleT) || (sn != kNibbleT && kn < sn);
          done = true;
        }
      } else {  // value in a branch slot: the key ends here
        st.kind = STOP_VALUE_UPDATE;
        st.stop = parent;
        st.attach = parent;
        st.slot = (uint8_t)pslot;
        st.base = (uint16_t)pos;
        done = true;
      }
    }
    if (want_mark) ds.mark(P, mark_ref);
  }
  if (active) {
    if (!done) raise_error(P.cnt, ERR_INTERNAL, 10);
    a.stop[i] = st;
  }
  dirty_flush(P, ds);
}

__device__ __forceinline__ bool same_group(const Stop &x, uint32_t tx, const Stop &y, uint32_t ty) {
  if (tx != ty) return false;
  if (x.kind == STOP_VALUE_UPDATE || y.kind == STOP_VALUE_UPDATE) return false;
  bool xs = x.kind == STOP_SHORT_SPLIT || x.kind == STOP_LEAF_UPDATE;
  bool ys = y.kind == STOP_SHORT_SPLIT || y.kind == STOP_LEAF_UPDATE;
  if (xs || ys) return xs && ys && x.stop == y.stop;
  if (x.kind != y.kind) return false;
  if (x.kind == STOP_ROOT_EMPTY) return true;
  return x.attach == y.attach && x.slot == y.slot;  // branch slot
}

__global__ __launch_bounds__(256) void k_group_flags(NuArgs a) {
  uint32_t i = grid_tid();
  uint32_t m = a.scal[SC_M_UNIQ];
  if (i >= a.n) return;
  if (i >= m) {
    a.ecount[i] = 0;
    return;
  }
  bool start = i == 0 || !same_group(a.stop[i - 1], a.u_trie[i - 1], a.stop[i], a.u_trie[i]);
  a.ecount[i] = start ? 1u : 0u;
}

// grp[] = exclusive scan of the start flags in ecount[]; make it inclusive,
// i.e. 1-based group ids.
__global__ __launch_bounds__(256) void k_group_fix(NuArgs a) {
  uint32_t i = grid_tid();
  if (i >= a.n) return;
  a.grp[i] += a.ecount[i];
}

// grp[] now holds the inclusive scan; group id = grp - 1.
__global__ __launch_bounds__(256) void k_group_info(NuArgs a) {
  uint32_t i = grid_tid();
  uint32_t m = a.scal[SC_M_UNIQ];
  if (i >= m) return;
  uint32_t g = a.grp[i] - 1;
  const Stop st = a.stop[i];
  bool start = i == 0 || a.grp[i - 1] != a.grp[i];
  if (start) {
    GroupInfo gi;
    gi.attach = st.attach;
    gi.split = (st.kind == STOP_SHORT_SPLIT || st.kind == STOP_LEAF_UPDATE) ? st.stop : REF_NULL;
    gi.trie = a.u_trie[i];
    gi.base = st.base;
    gi.slot = st.slot;
    gi.has_split = 0;
    a.ginfo[g] = gi;
  }
  if (i + 1 == m) a.scal[SC_NGROUP] = g + 1;
}

__global__ __launch_bounds__(256) void k_group_split(NuArgs a) {
  uint32_t i = grid_tid();
  uint32_t m = a.scal[SC_M_UNIQ];
  if (i >= m) return;
  if (a.stop[i].kind == STOP_SHORT_SPLIT) a.ginfo[a.grp[i] - 1].has_split = 1;
}


// A split group whose short node S is an extension can be interrupted in the
// sorted key order by keys that pass through S (they sort between the keys
// ordered before S's subtree and those after it).  The run after the
// subtree looks up the run before it (binary search for the start of the
// subtree's key block) and records that it belongs to that run's group.
__device__ bool key_before_block(const NuArgs &a, uint32_t j, uint32_t trie, const KeyRef &kr, const Short &S,
                                 uint32_t blen) {
  uint32_t tj = a.u_trie[j];
  if (tj != trie) return tj < trie;
  KeyRef kj = key_of(a, a.u_idx[j]);
  for (uint32_t q = 0; q < blen; q++) {
    uint32_t p = q < S.depth ? kr.nib(q) : short_nib(a.P, S, q);
    uint32_t x = q < kj.len ? kj.nib(q) : kNibbleT;
    if (x != p) return x == kNibbleT || x < p;  // terminator orders first
  }
  return false;  // kj has the whole prefix: inside the block
}

__global__ __launch_bounds__(256) void k_group_merge(NuArgs a) {
  uint32_t i = grid_tid();
  uint32_t m = a.scal[SC_M_UNIQ];
  if (i >= m) return;
  uint32_t run = a.grp[i] - 1;
  bool start = i == 0 || a.grp[i - 1] != a.grp[i];
  if (!start) return;
  a.merge[run] = run;
  const Stop st = a.stop[i];
  if (st.kind != STOP_SHORT_SPLIT || st.less) return;
  const Short &S = a.P.shrt[ref_idx(st.stop)];
  if (S.term) return;  // a leaf's subtree is one key, already grouped with it
  uint32_t trie = a.u_trie[i];
  KeyRef kr = key_of(a, a.u_idx[i]);
  uint32_t blen = S.depth + S.key_len;
  uint32_t lo = 0, hi = i;  // first j in [0, i) not before the block
  while (lo < hi) {
    uint32_t mid = (lo + hi) / 2;
    if (key_before_block(a, mid, trie, kr, S, blen))
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo == 0) return;
  uint32_t j = lo - 1;
  const Stop sj = a.stop[j];
  if (a.u_trie[j] == trie && sj.kind == STOP_SHORT_SPLIT && sj.stop == st.stop && sj.less) {
    a.merge[run] = a.grp[j] - 1;
    a.scal[SC_MERGE] = 1;
  }
}

// Every trie of the forest empty (a bulk build): each unique key stops at its
// trie's empty root, the groups are the tries, no key is an update and none
// splits a node, so each unique key is exactly one element.  One pass writes
// what k_descend, the group kernels and the element emission would: stops,
// 1-based group ids, group descriptors, merge targets, the element sequence
// and levels 0 .. 2 of the minimum tree.  Multi-trie batches give it the exclusive
// scan of the trie-start flags in grp (k_empty_flags); a single trie has one
// group.
__global__ __launch_bounds__(256) void k_empty_flags(NuArgs a) {
  const uint32_t i = grid_tid();
  const uint32_t m = a.scal[SC_M_UNIQ];
  if (i >= a.n) return;
  a.ecount[i] = i < m && (i == 0 || a.u_trie[i] != a.u_trie[i - 1]) ? 1u : 0u;
}

// Levels 1 and 2 of the split points' minimum tree (k_mt_level) over the
// block's 256 split points sp (255 past the last, above every valid value):
// a level-1 entry per 16 lanes and the block's level-2 entry.  Every thread
// of the block calls it.
__device__ __forceinline__ void mt_block_levels(const NuArgs &a, uint32_t sp) {
  __shared__ uint32_t wmin[4];
  uint32_t v = sp;
#pragma unroll
  for (int o = 1; o < 16; o <<= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o));
  const size_t off1 = mt_level_len(a.st_stride, 0);
  if (a.st_levels > 1 && (threadIdx.x & 15) == 0) a.st[off1 + (size_t)blockIdx.x * 16 + threadIdx.x / 16] = (uint8_t)v;
  v = min(v, (uint32_t)__shfl_xor((int)v, 16));
  v = min(v, (uint32_t)__shfl_xor((int)v, 32));
  if (__lane_id() == 0) wmin[threadIdx.x >> 6] = v;
  __syncthreads();
  if (a.st_levels > 2 && threadIdx.x == 0)
    a.st[off1 + mt_level_len(a.st_stride, 1) + blockIdx.x] = (uint8_t)min(min(wmin[0], wmin[1]), min(wmin[2], wmin[3]));
}

// Element i of a build into empty tries: its group and element entries;
// returns split point i (between elements i and i + 1: 0 where the next
// element starts another trie), or 255 for the last element.  (The stops,
// group numbers and merge targets the descent writes for other builds are
// read by no later kernel of this one; a single trie's element entries are
// the identity, e_ident.)
__device__ __forceinline__ uint32_t empty_element(const NuArgs &a, uint32_t i, uint32_t m, int multi) {
  const uint32_t t = a.u_trie[i];
  if (t >= a.ntries) raise_error(a.P.cnt, ERR_BAD_KEY, a.u_idx[i]);
  const bool start = i == 0 || (multi && a.u_trie[i - 1] != t);
  const uint32_t g = multi ? a.grp[i] + (start ? 1u : 0u) : 1u;  // 1-based
  if (start) {
    GroupInfo gi;
    gi.attach = REF_NULL;
    gi.split = REF_NULL;
    gi.trie = t;
    gi.base = 0;
    gi.slot = 0;
    gi.has_split = 0;
    a.ginfo[g - 1] = gi;
  }
  if (!a.e_ident) {
    a.e_src[i] = i;
    a.e_grp[i] = g - 1;
  }
  if (i + 1 == m) {
    a.scal[SC_NGROUP] = g;
    a.scal[SC_NE] = m;
    return 255;
  }
  const uint32_t sp = multi && a.u_trie[i + 1] != t ? 0u : (uint32_t)a.u_lcp[i + 1] + 1;
  a.st[i] = (uint8_t)sp;
  return sp;
}

__global__ __launch_bounds__(256) void k_empty_groups(NuArgs a, int multi) {
  const uint32_t i = grid_tid();
  const uint32_t m = a.scal[SC_M_UNIQ];
  mt_block_levels(a, i < m ? empty_element(a, i, m, multi) : 255u);
}

// Levels k0 .. of the minimum tree in one workgroup, each from the one below.
__global__ __launch_bounds__(256) void k_mt_top(NuArgs a, int k0) {
  const uint32_t ne = a.scal[SC_NE];
  uint32_t cnt = ne > 0 ? ne - 1 : 0;  // valid entries of level k - 1
  size_t off = 0;                      // and its offset
  for (int q = 0; q < k0 - 1; q++) {
    off += mt_level_len(a.st_stride, q);
    cnt = (cnt + 15) / 16;
  }
  for (int k = k0; k < a.st_levels; k++) {
    const size_t len = mt_level_len(a.st_stride, k - 1);
    const uint32_t nout = (cnt + 15) / 16;
    for (uint32_t b = threadIdx.x; b < nout; b += blockDim.x) {
      // the entry's 16 children in one load (levels are 16-byte multiples
      // at 16-byte aligned offsets); children past cnt do not count
      const uint4 v = *(const uint4 *)(a.st + off + 16 * (size_t)b);
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
      uint32_t mn = 255;
#pragma unroll
      for (uint32_t i = 0; i < 16; i++)
        if (16 * b + i < cnt) mn = min(mn, (w[i >> 2] >> (8 * (i & 3))) & 255u);
      a.st[off + len + b] = (uint8_t)mn;
    }
    __syncthreads();  // the level is complete for the next one
    off += len;
    cnt = nout;
  }
}

__global__ __launch_bounds__(256) void k_gkey(NuArgs a, uint32_t m) {
  uint32_t i = grid_tid();
  if (i >= m) return;
  uint32_t run = a.grp[i] - 1;
  uint32_t tgt = a.merge[run];
  // the target of a target is itself (a run joins at most one earlier run)
  a.gkey0[i] = tgt;
  a.perm0[i] = i;
}

__global__ __launch_bounds__(256) void k_permute(NuArgs a, const uint32_t *perm, uint32_t m) {
  uint32_t i = grid_tid();
  if (i >= m) return;
  uint32_t p = perm[i];
  a.u_idx2[i] = a.u_idx[p];
  a.u_kref2[i] = a.u_kref[p];
  if (a.u_win_ok) a.u_win2[i] = a.u_win[p];
  a.u_trie2[i] = a.u_trie[p];
  a.u_lcp2[i] = a.u_lcp[p];
  a.stop2[i] = a.stop[p];
}

__device__ __forceinline__ bool is_update(uint8_t kind) {
  return kind == STOP_LEAF_UPDATE || kind == STOP_VALUE_UPDATE;
}

// E layout per group: [keys ordered before the split node's subtree]
// [the split node as one virtual element] [keys after it]; updates emit nothing.
struct Emit {
  bool self, vbefore, vafter;
};

__device__ __forceinline__ Emit emit_of(const NuArgs &a, uint32_t i, uint32_t m) {
  Emit e{false, false, false};
  const Stop st = a.stop[i];
  if (is_update(st.kind)) return e;
  e.self = true;
  uint32_t g = a.grp[i];
  const GroupInfo &gi = a.ginfo[g - 1];
  if (!gi.has_split) return e;
  if (!st.less) {
    // first key ordered after the split subtree?
    bool first = true;
    if (i > 0 && a.grp[i - 1] == g) {
      const Stop p = a.stop[i - 1];
      first = is_update(p.kind) || p.less;
    }
    e.vbefore = first;
  } else {
    // last key ordered before the subtree, with no key after it in the group?
    bool last = true;
    uint32_t j = i + 1;
    if (j < m && a.grp[j] == g && is_update(a.stop[j].kind)) j++;
    if (j < m && a.grp[j] == g) last = false;
    e.vafter = last;
  }
  return e;
}

__global__ __launch_bounds__(256) void k_emit_count(NuArgs a) {
  uint32_t i = grid_tid();
  uint32_t m = a.scal[SC_M_UNIQ];
  if (i >= a.n) return;
  uint32_t c = 0;
  if (i < m) {
    Emit e = emit_of(a, i, m);
    c = (uint32_t)e.self + (uint32_t)e.vbefore + (uint32_t)e.vafter;
  }
  a.ecount[i] = c;
}

// ecount[] holds the exclusive scan.  Level 0 of the minimum tree gets, at
// split point q (between E[q] and E[q+1]), lcp(E[q], E[q+1]) + 1, or 0 across
// a group boundary.
__global__ __launch_bounds__(256) void k_emit_write(NuArgs a) {
  uint32_t i = grid_tid();
  uint32_t m = a.scal[SC_M_UNIQ];
  if (i >= m) return;
  Emit e = emit_of(a, i, m);
  if (!e.self) {
    if (i + 1 == m) a.scal[SC_NE] = a.ecount[i];
    return;
  }
  const Stop st = a.stop[i];
  uint32_t g = a.grp[i];
  uint32_t pos = a.ecount[i];
  uint8_t *sp = a.st;  // level 0
  const GroupInfo &gi = a.ginfo[g - 1];
  // does an earlier E element of this group exist, and what is its lcp to us?
  auto prev_in_group = [&](uint32_t ii) -> int {
    // index of the previous non-update unique key in the group, or -1
    if (ii == 0) return -1;
    uint32_t j = ii - 1;
    if (a.grp[j] != g) return -1;
    if (is_update(a.stop[j].kind)) {
      if (j == 0) return -1;
      j--;
      if (a.grp[j] != g) return -1;
    }
    return (int)j;
  };
  if (e.vbefore) {
    int pj = prev_in_group(i);
    a.e_src[pos] = 0x80000000u | ref_idx(gi.split);
    a.e_grp[pos] = g - 1;
    if (pos > 0) sp[pos - 1] = pj >= 0 ? (uint8_t)(a.stop[pj].div + 1) : 0;
    pos++;
    a.e_src[pos] = i;
    a.e_grp[pos] = g - 1;
    sp[pos - 1] = (uint8_t)(st.div + 1);
    pos++;
  } else {
    a.e_src[pos] = i;
    a.e_grp[pos] = g - 1;
    int pj = prev_in_group(i);
    if (pos > 0) sp[pos - 1] = pj >= 0 ? (uint8_t)(a.u_lcp[i] + 1) : 0;
    pos++;
  }
  if (e.vafter) {
    a.e_src[pos] = 0x80000000u | ref_idx(gi.split);
    a.e_grp[pos] = g - 1;
    sp[pos - 1] = (uint8_t)(st.div + 1);
    pos++;
  }
  if (i + 1 == m) a.scal[SC_NE] = pos;
}

// ------------------------------------------------------------- 4. build --

// Split-point LCP minima (SP[q] = lcp + 1 of split point q, 0 between
// groups) as a 16-ary minimum tree: entry b of level k is the minimum of
// entries 16b .. 16b + 15 of level k - 1.  A nearest-smaller query reads one
// 16-byte block per level it climbs and descends: for uniform keys the
// answer is almost always in the query's own block or the next one up,
// instead of one dependent byte load per level of a sparse table (21 levels
// for 2M split points).
constexpr int kMtMaxLevels = 8;  // 16^8 split points

// Level k from level k - 1, one entry per thread (entries past the level's
// valid count are never read by a query).
__global__ __launch_bounds__(256) void k_mt_level(NuArgs a, int k) {
  const uint32_t b = grid_tid();
  const uint32_t ne = a.scal[SC_NE];
  uint32_t cnt = ne > 0 ? ne - 1 : 0;  // valid entries of level k - 1
  size_t off = 0;
  for (int q = 0; q < k - 1; q++) {
    off += mt_level_len(a.st_stride, q);
    cnt = (cnt + 15) / 16;
  }
  if ((size_t)16 * b >= cnt) return;
  const uint8_t *lo = a.st + off + (size_t)16 * b;
  const uint4 v = *(const uint4 *)lo;
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
  uint32_t mn = 255;
#pragma unroll
  for (uint32_t i = 0; i < 16; i++)
    if (16 * b + i < cnt) mn = min(mn, (w[i >> 2] >> (8 * (i & 3))) & 255u);
  a.st[off + mt_level_len(a.st_stride, k - 1) + b] = (uint8_t)mn;
}

// The tree's levels are located from its level-0 length (stride) alone,
// level k's byte offset by a short loop and its valid entries as
// ceil(nsp / 16^k): tables indexed by a level known only at run time would
// live in scratch memory.
struct ST {
  const uint8_t *t;
  uint32_t stride;  // bytes of level 0 (mt_level_len geometry)
  int levels;
  uint32_t nsp;
  __device__ __forceinline__ uint32_t off(int k) const {  // byte offset of level k
    uint32_t o = 0, len = stride;
    for (int q = 0; q < k; q++) {
      o += len;
      len = ((len / 16 + 15) / 16) * 16;
    }
    return o;
  }
  __device__ __forceinline__ uint32_t cnt(int k) const {  // valid entries of level k
    return (uint32_t)(((uint64_t)nsp + (1ull << (4 * k)) - 1) >> (4 * k));
  }
  __device__ __forceinline__ uint8_t at(int k, uint32_t q) const { return t[off(k) + q]; }
  // 16-bit mask of the valid entries of the block of level k at base
  // (a multiple of 16) whose value is below lim (lim <= 256)
  __device__ __forceinline__ uint32_t below(int k, uint32_t base, uint32_t lim) const {
    const uint4 v = *(const uint4 *)(t + off(k) + base);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    uint32_t m = 0;
#pragma unroll
    for (uint32_t i = 0; i < 16; i++) m |= (uint32_t)(((w[i >> 2] >> (8 * (i & 3))) & 255u) < lim) << i;
    const uint32_t valid = cnt(k) - base;  // > 0 for every block a query reads
    return valid >= 16 ? m : m & ((1u << valid) - 1);
  }
  // rightmost (leftmost) entry below lim under level-k entry p, at level 0
  // (the level loops stay rolled: the code is shared by many call sites)
  __device__ __forceinline__ uint32_t down_last(int k, uint32_t p, uint32_t lim) const {
#pragma unroll 1
    for (; k > 0; k--) {
      const uint32_t base = 16 * p;
      p = base + 31 - __clz(below(k - 1, base, lim));
    }
    return p;
  }
  __device__ __forceinline__ uint32_t down_first(int k, uint32_t p, uint32_t lim) const {
#pragma unroll 1
    for (; k > 0; k--) {
      const uint32_t base = 16 * p;
      p = base + __ffs(below(k - 1, base, lim)) - 1;
    }
    return p;
  }
  // largest q <= j0 with SP[q] < lim, or -1 (j0 may be -1)
  __device__ __forceinline__ int last_below(int64_t j0, uint32_t lim) const {
    int64_t idx = j0;
#pragma unroll 1
    for (int k = 0; k < levels && idx >= 0; k++) {
      const uint32_t base = (uint32_t)idx & ~15u;
      const uint32_t m = below(k, base, lim) & (0xffffu >> (15 - ((uint32_t)idx - base)));
      if (m) return (int)down_last(k, base + 31 - __clz(m), lim);
      idx = (int64_t)(base / 16) - 1;
    }
    return -1;
  }
  // smallest q >= j0 with SP[q] < lim, or nsp
  __device__ __forceinline__ uint32_t first_below(uint32_t j0, uint32_t lim) const {
    uint32_t idx = j0;
#pragma unroll 1
    for (int k = 0; k < levels && idx < cnt(k); k++) {
      const uint32_t base = idx & ~15u;
      const uint32_t m = below(k, base, lim) & (0xffffu << (idx - base)) & 0xffffu;
      if (m) return down_first(k, base + __ffs(m) - 1, lim);
      idx = base / 16 + 1;
    }
    return nsp;
  }
  // largest q < j with SP[q] < L, or -1
  __device__ __forceinline__ int prev_less(uint32_t j, uint32_t L) const { return last_below((int64_t)j - 1, L); }
  // smallest q > j with SP[q] < L, or nsp
  __device__ __forceinline__ uint32_t next_less(uint32_t j, uint32_t L) const { return first_below(j + 1, L); }
  // smallest q >= j with SP[q] <= L, or nsp
  __device__ __forceinline__ uint32_t next_leq(uint32_t j, uint32_t L) const { return first_below(j, L + 1); }
};

__device__ __forceinline__ ST st_of(const NuArgs &a) {
  ST s;
  s.t = a.st;
  uint32_t ne = a.scal[SC_NE];
  s.nsp = ne > 0 ? ne - 1 : 0;
  s.levels = min(a.st_levels, kMtMaxLevels);
  s.stride = (uint32_t)a.st_stride;
  return s;
}

// A workgroup's window of the split points in LDS: level 0 entries
// [wb, wb + kSpWin) and their level-1 .. level-3 minima, queried through an
// ST over window-relative positions (its nsp = the window's valid length).
// With 4,096 entries each side of the tile, only a split point whose
// interval reaches further (an LCP of 2 or less for uniform keys) falls back
// to the global tree.
constexpr uint32_t kSpTile = 256;                      // split points per workgroup
constexpr uint32_t kSpHalo = 4096;                     // window entries on each side of the tile
constexpr uint32_t kSpWin = kSpHalo + kSpTile + kSpHalo;
constexpr uint32_t kSpL1 = kSpWin / 16;                         // level-1 entries (a multiple of 16)
constexpr uint32_t kSpL2 = (kSpL1 + 15) / 16;                   // level-2 entries
constexpr uint32_t kSpL2Pad = (kSpL2 + 15) / 16 * 16;
constexpr uint32_t kSpL3 = (kSpL2 + 15) / 16;                   // level-3 entries (<= 16)
constexpr uint32_t kSpWinBytes = kSpWin + kSpL1 + kSpL2Pad + 16;
static_assert(kSpWin % 256 == 0 && kSpL3 <= 16, "window geometry");
static_assert(kSpL1 == ((kSpWin / 16 + 15) / 16) * 16 && kSpL2Pad == ((kSpL1 / 16 + 15) / 16) * 16 &&
                  ((kSpL2Pad / 16 + 15) / 16) * 16 == 16,
              "the window's levels follow the tree's level lengths");

// Loads the window of the tile starting at t0 (a multiple of kSpTile) into
// w (kSpWinBytes, 16-byte aligned) and returns its ST.  Every thread of the
// block calls it (it synchronises the block).
__device__ __forceinline__ ST st_window(const ST &g, uint32_t t0, uint8_t *w, uint32_t &wb) {
  wb = t0 >= kSpHalo ? t0 - kSpHalo : 0u;
  const uint32_t n0 = g.nsp > wb ? min(kSpWin, g.nsp - wb) : 0u;
  uint8_t *w1 = w + kSpWin, *w2 = w1 + kSpL1, *w3 = w2 + kSpL2Pad;
  // level 0, 16 bytes per chunk (the tree's level 0 is padded to a multiple
  // of 256 bytes, so whole 16-byte reads stay inside it), and each chunk's
  // level-1 minimum over the valid entries; the thread's chunk loads are all
  // issued before any is used
  constexpr uint32_t kPer = (kSpL1 + kSpTile - 1) / kSpTile;
  uint4 v[kPer];
#pragma unroll
  for (uint32_t q = 0; q < kPer; q++) {
    const uint32_t c = threadIdx.x + kSpTile * q;
    v[q] = make_uint4(0, 0, 0, 0);
    if (c < kSpL1 && 16 * c < n0) v[q] = *(const uint4 *)(g.t + wb + 16 * c);
  }
#pragma unroll
  for (uint32_t q = 0; q < kPer; q++) {
    const uint32_t c = threadIdx.x + kSpTile * q;
    if (c >= kSpL1) break;
    uint32_t mn = 255;
    const uint32_t wd[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
#pragma unroll
    for (uint32_t i = 0; i < 16; i++)
      if (16 * c + i < n0) mn = min(mn, (wd[i >> 2] >> (8 * (i & 3))) & 255u);
    ((uint4 *)w)[c] = v[q];
    w1[c] = (uint8_t)mn;
  }
  __syncthreads();
  for (uint32_t c = threadIdx.x; c < kSpL2Pad; c += blockDim.x) {
    uint32_t mn = 255;
    if (c < kSpL2)
      for (uint32_t i = 0; i < 16; i++) mn = min(mn, (uint32_t)w1[16 * c + i]);
    w2[c] = (uint8_t)mn;
  }
  __syncthreads();
  if (threadIdx.x < 16) {
    uint32_t mn = 255;
    if (threadIdx.x < kSpL3)
      for (uint32_t i = 0; i < 16; i++) mn = min(mn, (uint32_t)w2[16 * threadIdx.x + i]);
    w3[threadIdx.x] = (uint8_t)mn;
  }
  __syncthreads();
  ST s;  // the window's levels follow the tree's geometry with a level 0 of kSpWin bytes
  s.t = w;
  s.levels = 4;
  s.nsp = n0;
  s.stride = kSpWin;
  return s;
}

// Per split point: find its branch (the representative is the leftmost split
// point with the same LCP inside the maximal interval whose LCPs are >= it).
// bcnt[block] = the block's representatives, from which k_build_links
// ranks them.  A representative also records its interval's bounding split
// points (plohi).  The queries run on the workgroup's window of split points
// in LDS.
__global__ __launch_bounds__(256) void k_split_reps(NuArgs a, uint32_t ne_max) {
  // the sort needs a fallback pass, or an error is pending: the build waits
  // for the host (which may run it before reading either)
  if (a.P.cnt->flag_fallback || a.P.cnt->err) return;
  static_assert(kSpTile == 256, "one split point per thread");
  __shared__ __align__(16) uint8_t win[kSpWinBytes];
  const uint32_t j = blockIdx.x * kSpTile + threadIdx.x;
  // the build's dirty-list base: the list does not change until the build's
  // last kernel advances it
  if (j == 0) a.scal[SC_DBASE] = a.P.cnt->n_dirty;
  const ST g = st_of(a);
  uint32_t wb;
  const ST w = st_window(g, blockIdx.x * kSpTile, win, wb);
  bool rep = false;
  if (j < g.nsp) {
    const uint32_t jw = j - wb;  // inside the window: j < wb + nsp of the window
    const uint32_t L = w.at(0, jw);
    uint32_t r = ~0u;
    if (L != 0) {
      // nearest smaller to the left: in the window, else (window not at the
      // start) the global tree
      int lo = w.prev_less(jw, L);
      lo = lo >= 0 ? (int)wb + lo : wb == 0 ? -1 : g.prev_less(j, L);
      // the first entry <= L after it: at or before j, so in the window
      // unless lo is left of it
      r = (int64_t)lo + 1 >= (int64_t)wb ? wb + w.next_leq((uint32_t)(lo + 1) - wb, L)
                                         : g.next_leq((uint32_t)(lo + 1), L);
      rep = r == j;
      if (rep) {
        uint32_t hi = w.next_less(jw, L);
        hi = hi < w.nsp ? wb + hi : wb + w.nsp >= g.nsp ? g.nsp : g.next_less(j, L);
        a.plohi[j] = (uint64_t)(uint32_t)lo | (uint64_t)hi << 32;
      }
    }
    a.repof[j] = r;
  }
  const uint32_t reps = (uint32_t)__syncthreads_count(rep);
  if (threadIdx.x == 0) a.bcnt[blockIdx.x] = reps;
}

// Node slots of this build, all fixed by position (no allocator atomics):
//   branch of representative rank r   : base_branch + r
//   leaf / remainder short of element e: base_short + e
//   extension short of branch rank r   : base_short + NE + r
//   value of new-key element e         : base_value + e
//   (an updated key's value node is rewritten in place)
// Dirty-list entries: dbase + e (element short), dbase + NE + r (branch),
// dbase + NE + R + r (extension); REF_NULL where no node was made.
struct Slots {
  uint32_t ne, nrep, dbase;
};
__device__ __forceinline__ void put_dirty(const NuArgs &a, uint32_t slot, ref_t r) {
  if (slot < a.P.cap_dirty)
    a.P.dirty[slot] = r;
  else
    raise_error(a.P.cnt, ERR_POOL, 4);
}

// One more child linked into the build's branch of rank `rank` (the record
// the classification bins it by); a child in slot 16 makes it read the node.
__device__ __forceinline__ void count_child(const NuArgs &a, uint32_t rank, uint32_t slot) {
  if (slot == kNibbleT)
    atomicOr(&a.bdesc[rank], kBranchRecChase);
  else
    atomicAdd(&a.bdesc[rank], kBranchRecChild);
}

// Source of an E element's key nibbles.
struct ElemKey {
  uint64_t off;       // key arena offset of the source key
  uint32_t start;     // source nibble index of absolute position `origin`
  uint32_t origin;
  uint32_t end;       // absolute end of the element's path (exclusive)
  uint8_t term;       // path ends with the terminator
  bool virt;
  bool has_win;       // win holds nibbles W .. W + 15 of the key
  uint32_t uidx;      // unique index (new key) or short node idx (virtual)
  uint64_t win;
};

__device__ __forceinline__ ElemKey elem_key(const NuArgs &a, uint32_t e) {
  ElemKey k;
  uint32_t src = a.e_ident ? e : a.e_src[e];
  if (src & 0x80000000u) {
    const Short &S = a.P.shrt[src & 0x7fffffffu];
    k.off = S.key_off;
    k.start = S.key_start;
    k.origin = S.depth;
    k.end = S.depth + S.key_len;
    k.term = S.term;
    k.virt = true;
    k.has_win = false;
    k.win = 0;
    k.uidx = src & 0x7fffffffu;
  } else {
    const uint64_t r = a.u_kref[src];
    k.off = a.kbase + (r & kKrefOff);
    k.start = 0;
    k.origin = 0;
    k.end = (uint32_t)(r >> 48);
    k.term = 1;
    k.virt = false;
    k.has_win = a.u_win_ok != 0;
    k.win = k.has_win ? a.u_win[src] : 0ull;
    k.uidx = src;
  }
  return k;
}

// Nibble j of the element's path; W is the batch's common prefix (the sort
// window's first nibble): nibbles W .. W + 15 of a new key come from its
// window when the window is at hand.
__device__ __forceinline__ uint32_t elem_nib(const Pools &P, const ElemKey &k, uint32_t j, uint32_t W) {
  if (k.term && j + 1 == k.end) return kNibbleT;
  if (k.has_win && j >= W && j - W < 16) return (uint32_t)(k.win >> (60 - 4 * (j - W))) & 15u;
  return packed_nibble(P.keys + k.off, k.start + (j - k.origin));
}

__device__ __forceinline__ void init_short(Short &S, uint64_t key_off, uint32_t key_start, uint32_t key_len,
                                           uint32_t depth, uint8_t term, ref_t child) {
  uint64_t *w = (uint64_t *)&S;
#pragma unroll
  for (int q = 0; q < 8; q++) w[q] = 0;
  S.key_off = key_off;
  S.child = child;
  S.key_start = (uint16_t)key_start;
  S.key_len = (uint16_t)key_len;
  S.depth = (uint16_t)depth;
  S.term = term;
  S.dirty = 1;
}

__device__ __forceinline__ void install(const NuArgs &a, const GroupInfo &gi, ref_t r) {
  if (gi.attach == REF_NULL)
    a.P.tries[gi.trie].root = r;
  else
    a.P.branch[ref_idx(gi.attach)].child[gi.slot] = r;
}

// Branches: link each to its parent (through an extension when depths skip).
template <class Sink>
__device__ __forceinline__ void link_branch(const NuArgs &a, uint32_t j, uint32_t rank, const Slots &q,
                                            const Sink &sink) {
  ST s = st_of(a);
  if (j >= s.nsp) return;
  uint32_t L = s.at(0, j);
  if (L == 0 || a.repof[j] != j) return;
  // the interval's bounding split points, found by k_split_reps
  const uint64_t lh = a.plohi[j];
  const int lo = (int)(uint32_t)lh;
  const uint32_t hi = (uint32_t)(lh >> 32);
  uint32_t La = lo >= 0 ? s.at(0, (uint32_t)lo) : 0u;
  uint32_t Lb = hi < s.nsp ? s.at(0, hi) : 0u;
  uint32_t e0 = (uint32_t)(lo + 1);
  ElemKey ek = elem_key(a, e0);
  uint32_t depth = L - 1, lo_depth, parent = 0, slot = 0;
  bool is_root = La == 0 && Lb == 0;
  GroupInfo gi;
  if (is_root) {
    gi = a.ginfo[a.e_ident ? 0u : a.e_grp[e0]];
    lo_depth = gi.base;
  } else {
    uint32_t p = La >= Lb ? (uint32_t)lo : hi;
    uint32_t Lp = max(La, Lb) - 1;
    parent = a.repof[p];
    slot = elem_nib(a.P, ek, Lp, a.scal[SC_W]);
    lo_depth = Lp + 1;
  }
  ref_t me = make_ref(T_BRANCH, a.base_branch + rank);
  ref_t ext = REF_NULL;
  if (depth > lo_depth) {
    uint32_t sid = a.base_short + q.ne + rank;
    if (sid >= a.P.cap_short) {
      raise_error(a.P.cnt, ERR_POOL, 2);
    } else {
      init_short(a.P.shrt[sid], ek.off, ek.start + (lo_depth - ek.origin), depth - lo_depth, lo_depth, 0, me);
      ext = make_ref(T_SHORT, sid);
      me = ext;
    }
  }
  put_dirty(a, q.dbase + q.ne + q.nrep + rank, ext);
  if (is_root)
    install(a, gi, me);
  else
    sink(a, parent, slot, me);
}

// Elements: new keys become a leaf short node + value node; a split short
// node's remainder is re-attached below the new branch.
__device__ __forceinline__ void init_value(const NuArgs &a, uint32_t vid, uint32_t idx, uint64_t *addr = nullptr,
                                           uint32_t *size = nullptr) {
  Value &V = a.P.value[vid];
  int64_t vs = a.vidx[2 * idx], ve = a.vidx[2 * idx + 1];
  V.addr = a.vbase + (uint64_t)vs;
  V.hp = a.hps ? a.hps[idx] : 0ull;
  V.size = (uint32_t)(ve - vs + 1);
  V.pad = 0;
  V.seq = 0;
  if (addr) *addr = V.addr;
  if (size) *size = V.size;
}

template <class Sink>
__device__ __forceinline__ void link_elem(const NuArgs &a, uint32_t e, const Slots &q, const Sink &sink) {
  ST s = st_of(a);
  if (e == 0) {
    // the build's node and dirty-list slots are all assigned by position:
    // advance the pool counters past them (no kernel of the build reads them)
    a.P.cnt->n_branch = a.base_branch + q.nrep;
    a.P.cnt->n_short = a.base_short + q.ne + q.nrep;
    a.P.cnt->n_value = a.base_value + q.ne;
    a.P.cnt->n_dirty = q.dbase + q.ne + 2 * q.nrep;
  }
  if (e >= q.ne) return;
  ElemKey ek = elem_key(a, e);
  uint32_t La = e > 0 ? s.at(0, e - 1) : 0u;
  uint32_t Lb = e < s.nsp ? s.at(0, e) : 0u;
  uint32_t start, slot = 0, parent = 0;
  bool is_root = La == 0 && Lb == 0;
  GroupInfo gi;
  if (is_root) {
    gi = a.ginfo[a.e_ident ? 0u : a.e_grp[e]];
    start = gi.base;
  } else {
    uint32_t p = La >= Lb ? e - 1 : e;
    uint32_t Lp = max(La, Lb) - 1;
    parent = a.repof[p];
    slot = elem_nib(a.P, ek, Lp, a.scal[SC_W]);
    start = Lp + 1;
  }
  ref_t child;
  uint64_t va = 0;
  uint32_t vsz = 0;
  if (ek.virt) {
    child = a.P.shrt[ek.uidx].child;
  } else {
    uint32_t vid = a.base_value + e;
    if (vid >= a.P.cap_value) {
      raise_error(a.P.cnt, ERR_POOL, 3);
      a.leaf_pk[e] = kLeafPackedNone;
      return;
    }
    init_value(a, vid, a.u_idx[ek.uidx], &va, &vsz);
    child = make_ref(T_VALUE, vid);
  }
  ref_t me = child, sh = REF_NULL;
  uint32_t pk = kLeafPackedNone;
  if (start < ek.end) {
    uint32_t sid = a.base_short + e;
    if (sid >= a.P.cap_short) {
      raise_error(a.P.cnt, ERR_POOL, 2);
    } else {
      const uint32_t ks = ek.start + (start - ek.origin), kl = ek.end - start;
      init_short(a.P.shrt[sid], ek.off, ks, kl, start, ek.term, child);
      sh = make_ref(T_SHORT, sid);
      me = sh;
      if (!ek.virt) {  // a new leaf: record its hashing bin and metadata
        pk = leaf_packed(kl - ek.term, vsz);
        leaf_meta_words(ek.off, ks, kl - ek.term, ek.term, vsz, va, a.leaf_meta + 2 * (size_t)e);
      }
    }
  }
  a.leaf_pk[e] = pk;
  put_dirty(a, q.dbase + e, sh);
  if (is_root)
    install(a, gi, me);
  else
    sink(a, parent, slot, me);
}

// The build's branches and links in one launch.  Block b ranks the
// representatives among its split points [256b, 256b + 256) (the block's
// offset from k_split_reps' counts, then a block scan: rank r is branch
// base_branch + r, bid records it, SC_NREP_TOT the total) and owns their
// branches: it builds their child arrays in LDS from every link whose parent
// it owns -- most links, since a branch's children sit beside its
// representative in batch order -- and then writes each owned branch whole,
// once, with coalesced stores.  A link whose parent another block owns goes
// to the cross list (xp, xs, xm), which k_link_cross applies afterwards; no
// other branch line is written piecewise.  Thread j links the branch whose
// representative is split point j, if any, and element j.
constexpr uint32_t kCrossPerBlock = 2 * kSpTile;  // a thread makes at most two links
struct OwnedLinks {
  ref_t (*kid)[17];
  uint32_t *krec;
  const uint16_t *own;
  uint32_t *xn;  // the block's cross links so far (LDS)
  uint32_t lo;   // the block's first split point
  __device__ __forceinline__ void operator()(const NuArgs &a, uint32_t parent, uint32_t slot, ref_t me) const {
    if (parent - lo < kSpTile) {  // a representative of this block
      const uint32_t l = own[parent - lo];
      kid[l][slot] = me;
      if (slot == kNibbleT)
        atomicOr(&krec[l], kBranchRecChase);
      else
        atomicAdd(&krec[l], kBranchRecChild);
    } else {  // into the block's own segment of the cross list
      const uint32_t x = blockIdx.x * kCrossPerBlock + atomicAdd(xn, 1u);
      a.xp[x] = parent;
      a.xs[x] = (uint8_t)slot;
      a.xm[x] = me;
    }
  }
};

__global__ __launch_bounds__(256) void k_build_links(NuArgs a, int pre) {
  // the sort needs a fallback pass, or an error is pending: the build waits
  // for the host (which may run it before reading either)
  if (a.P.cnt->flag_fallback || a.P.cnt->err) return;
  static_assert(kSpTile == 256, "one split point per thread");
  static_assert(offsetof(Branch, hash_size) == 32 && offsetof(Branch, depth) == 34 && offsetof(Branch, dirty) == 36 &&
                    offsetof(Branch, lock) == 40 && offsetof(Branch, child) == 48 && sizeof(Branch) == 128,
                "owned branches are written as eight 16-byte words");
  __shared__ uint32_t sh[4];
  __shared__ uint64_t sh2[4];
  __shared__ ref_t kid[kSpTile][17];
  __shared__ uint32_t krec[kSpTile];   // the owned branches' records (bdesc): depth, children, slot-16 flag
  __shared__ uint16_t own[kSpTile];    // per split point of the block: its branch's local index
  __shared__ uint32_t xn;
  const uint32_t j = grid_tid();
  const ST s = st_of(a);
  const uint32_t L = j < s.nsp ? s.at(0, j) : 0u;
  const bool rep = L != 0 && a.repof[j] == j;
  // the block's offset among the representatives and their total
  uint32_t base, nrep;
  if (pre) {
    base = a.bcnt[blockIdx.x];
    nrep = a.scal[SC_NREP_TOT];  // the count scan's total
  } else {
    uint64_t part = 0;  // representatives before the block | all of them << 32
#pragma unroll
    for (uint32_t q = 0; q < kSelfSumBlocks / 256; q++) {
      const uint32_t t = threadIdx.x + 256 * q;
      const uint64_t c = t < gridDim.x ? a.bcnt[t] : 0u;
      part += (t < blockIdx.x ? c : 0ull) | c << 32;
    }
    uint64_t tot;
    block_exclusive_scan_256<uint64_t>(part, sh2, &tot);
    base = (uint32_t)tot;
    nrep = (uint32_t)(tot >> 32);
  }
  uint32_t btot;
  const uint32_t lrank = block_exclusive_scan_256<uint32_t>(rep, sh, &btot);
  const uint32_t rank = base + lrank;
  if (blockIdx.x + 1 == gridDim.x && threadIdx.x == 0) a.scal[SC_NREP_TOT] = nrep;
  Slots q;
  q.ne = a.scal[SC_NE];
  q.nrep = nrep;
  q.dbase = a.scal[SC_DBASE];
  own[threadIdx.x] = rep ? (uint16_t)lrank : (uint16_t)0xffffu;
  if (threadIdx.x == 0) xn = 0;
  for (uint32_t k = threadIdx.x; k < btot * 17; k += kSpTile) (&kid[0][0])[k] = REF_NULL;
  if (rep) {
    krec[lrank] = L - 1;
    a.bid[j] = rank;
    put_dirty(a, q.dbase + q.ne + rank, make_ref(T_BRANCH, a.base_branch + rank));
  }
  __syncthreads();  // child arrays cleared, local indices set
  const OwnedLinks sink{kid, krec, own, &xn, blockIdx.x * kSpTile};
  link_branch(a, j, rank, q, sink);
  link_elem(a, j, q, sink);
  __syncthreads();  // every owned link in LDS
  // the owned branches, 16 bytes per thread: hash (zero), hash_size 0 |
  // depth | dirty 1 | lock 0, child[0 .. 15], child[16] and padding
  const uint32_t b0 = a.base_branch + base;
  for (uint32_t k = threadIdx.x; k < btot * 8; k += kSpTile) {
    const uint32_t l = k / 8, w = k % 8;
    if (b0 + l >= a.P.cap_branch) {
      raise_error(a.P.cnt, ERR_POOL, 1);
      continue;
    }
    uint4 v = make_uint4(0, 0, 0, 0);
    if (w == 2) {
      v.x = (krec[l] & 0xffffu) << 16;  // the representative's L - 1: the depth
      v.y = 1;
    } else if (w >= 3 && w <= 6) {
      v = make_uint4(kid[l][4 * (w - 3)], kid[l][4 * (w - 3) + 1], kid[l][4 * (w - 3) + 2], kid[l][4 * (w - 3) + 3]);
    } else if (w == 7) {
      v.x = kid[l][16];
    }
    ((uint4 *)&a.P.branch[b0 + l])[w] = v;
  }
  for (uint32_t k = threadIdx.x; k < btot; k += kSpTile) a.bdesc[base + k] = krec[k];
  if (threadIdx.x == 0) a.xcnt[blockIdx.x] = xn;
}

// The cross links of k_build_links (block b's in its segment of the list):
// each into its parent's child slot, and counted in the parent's record.
__global__ __launch_bounds__(256) void k_link_cross(NuArgs a) {
  if (a.P.cnt->flag_fallback || a.P.cnt->err) return;
  const uint32_t n = a.xcnt[blockIdx.x];
  for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
    const uint32_t x = blockIdx.x * kCrossPerBlock + k;
    const uint32_t rank = a.bid[a.xp[x]], slot = a.xs[x];
    if (a.base_branch + rank >= a.P.cap_branch) continue;  // raised by the branch's own block
    a.P.branch[a.base_branch + rank].child[slot] = a.xm[x];
    count_child(a, rank, slot);
  }
}

// Existing keys: swap in a new value node (runs before the links, so a leaf
// that is also split carries the new value into its remainder).
// The key's value node is rewritten in place (no node is allocated per
// update; only this key refers to it).
__global__ __launch_bounds__(256) void k_updates(NuArgs a) {
  uint32_t i = grid_tid();
  uint32_t m = a.scal[SC_M_UNIQ];
  if (i >= m) return;
  Stop st = a.stop[i];
  if (!is_update(st.kind)) return;
  const ref_t vr = st.kind == STOP_LEAF_UPDATE ? a.P.shrt[ref_idx(st.stop)].child
                                               : a.P.branch[ref_idx(st.stop)].child[16];
  if (ref_type(vr) != T_VALUE) {
    raise_error(a.P.cnt, ERR_INTERNAL, 13);
    return;
  }
  init_value(a, ref_idx(vr), a.u_idx[i]);
}


}  // namespace

// ------------------------------------------------------------- host side --

namespace {
__global__ void k_init_scalars(uint32_t *scal) {
  uint32_t t = threadIdx.x;
  if (t < SC_COUNT) scal[t] = t == SC_W ? 0xffffffffu : 0u;
}
}  // namespace

hipError_t nu_init_scalars(const NuArgs &a, hipStream_t s) {
  k_init_scalars<<<1, 64, 0, s>>>(a.scal);
  return hipGetLastError();
}

hipError_t nu_prepare_keys(const NuArgs &a, void *temp, hipStream_t s) {
  // per-block partial results live in ecount / grp, unused until grouping
  k_keys_scan<<<nblk(a.n), kB, 0, s>>>(a, a.ecount, a.grp);
  GMPT_LAUNCH_CHECK();
  k_keys_reduce<<<1, kB, 0, s>>>(a, a.ecount, a.grp, nblk(a.n));
  GMPT_LAUNCH_CHECK();
  k_keys_pack<<<nblk(a.n), kB, 0, s>>>(a);
  return hipGetLastError();
}

namespace {

// After a stable sort by the windows' top kTopBits bits: every run of equal
// top bits is sorted by the whole window with a stable insertion sort by the
// thread at the run's start (runs of uniform keys hold one or two entries).
// A run longer than kFixMaxRun sets bit 1 of flag_fallback: the host then
// sorts by all 64 bits.
constexpr int kTopBits = 24;
constexpr uint32_t kFixMaxRun = 64;

__global__ __launch_bounds__(256) void k_fix_runs(NuArgs a, uint64_t *sk, uint32_t *sv) {
  uint32_t j = grid_tid();
  if (j >= a.n) return;
  const uint64_t top = sk[j] >> (64 - kTopBits);
  if (j > 0 && (sk[j - 1] >> (64 - kTopBits)) == top) return;  // not a run start
  uint32_t L = 1;
  while (j + L < a.n && L <= kFixMaxRun && (sk[j + L] >> (64 - kTopBits)) == top) L++;
  if (L == 1) return;
  if (L > kFixMaxRun) {
    atomicOr(&a.P.cnt->flag_fallback, 2u);
    return;
  }
  for (uint32_t q = 1; q < L; q++) {
    const uint64_t k = sk[j + q];
    const uint32_t v = sv[j + q];
    uint32_t p = q;
    while (p > 0 && sk[j + p - 1] > k) {
      sk[j + p] = sk[j + p - 1];
      sv[j + p] = sv[j + p - 1];
      p--;
    }
    sk[j + p] = k;
    sv[j + p] = v;
  }
}

// OR and AND of the windows per block (no global atomics: one slot pair per
// block), then of the block results.
__device__ __forceinline__ void or_and_block(uint64_t &o, uint64_t &n) {
  __shared__ uint64_t so[4], sn[4];
  for (int off = 32; off > 0; off >>= 1) {
    o |= (uint64_t)__shfl_xor((unsigned long long)o, off);
    n &= (uint64_t)__shfl_xor((unsigned long long)n, off);
  }
  if (__lane_id() == 0) {
    so[threadIdx.x >> 6] = o;
    sn[threadIdx.x >> 6] = n;
  }
  __syncthreads();
  o = so[0] | so[1] | so[2] | so[3];
  n = sn[0] & sn[1] & sn[2] & sn[3];
}
__global__ __launch_bounds__(256) void k_wmask_part(uint32_t n, const uint64_t *sk, uint64_t *part) {
  const uint32_t j = grid_tid();
  uint64_t o = 0, m = ~0ull;
  if (j < n) o = m = sk[j];
  or_and_block(o, m);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = o;
    part[2 * blockIdx.x + 1] = m;
  }
}
__global__ __launch_bounds__(256) void k_wmask_final(uint32_t nb, const uint64_t *part, uint32_t *scal) {
  uint64_t o = 0, m = ~0ull;
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
    o |= part[2 * b];
    m &= part[2 * b + 1];
  }
  or_and_block(o, m);
  if (threadIdx.x == 0) {
    *(uint64_t *)(scal + SC_WOR) = o;
    *(uint64_t *)(scal + SC_WAND) = m;
  }
}
// Keeps the window bits of vmask, in order, at the top of the key.
__global__ __launch_bounds__(256) void k_compress_windows(uint32_t n, uint64_t *sk, uint64_t vmask) {
  const uint32_t j = grid_tid();
  if (j >= n) return;
  sk[j] = gather_bits(sk[j], vmask) << (64 - __popcll(vmask));
}

}  // namespace

hipError_t nu_window_mask(const NuArgs &a, const uint64_t *sk, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  const unsigned nb = nblk(a.n);
  if (4ull * nb > a.n) return hipErrorInvalidValue;  // partials do not fit in ecount
  uint64_t *part = (uint64_t *)a.ecount;
  k_wmask_part<<<nb, kB, 0, s>>>(a.n, sk, part);
  GMPT_LAUNCH_CHECK();
  k_wmask_final<<<1, kB, 0, s>>>(nb, part, a.scal);
  return hipGetLastError();
}

hipError_t nu_sort(NuArgs &a, int full, uint32_t trie_bits, void *temp, hipStream_t s, uint64_t **sk,
                   uint32_t **sv, bool win_all_bits, int *cur_io, uint64_t vmask, uint64_t *sk_mask) {
  uint64_t *k[2] = {a.sk0, a.sk1};
  uint32_t *v[2] = {a.sv0, a.sv1};
  int cur = cur_io ? *cur_io : 0, which = 0;
  hipError_t e;
  if (!full && !win_all_bits && a.n <= kBucketSortMaxN) {
    // whole windows (ties by input position) in one pass plus a sort of each
    // top-8-bit bucket in LDS; a bucket too large for LDS asks for pass 1
    e = bucket_sort_u64(k[cur], v[cur], k[cur ^ 1], v[cur ^ 1], a.n, &which, temp, s, &a.P.cnt->flag_fallback);
    if (e != hipSuccess) return e;
    cur ^= which;
  } else if (!full && !win_all_bits) {
    // top bits, then each run of equal top bits by the whole window
    e = radix_sort_u64(k[cur], v[cur], k[cur ^ 1], v[cur ^ 1], a.n, 64 - kTopBits, 64, &which, temp, s);
    if (e != hipSuccess) return e;
    cur ^= which;
    k_fix_runs<<<nblk(a.n), kB, 0, s>>>(a, k[cur], v[cur]);
    GMPT_LAUNCH_CHECK();
  } else if (!full) {
    // compressed to the varying bits (all of them: no fix-up needed), or all 64
    const int pop = vmask ? __builtin_popcountll(vmask) : 64;
    const int lo = pop < 64 ? 64 - 8 * ((pop + 7) / 8) : 0;
    if (pop < 64) {
      k_compress_windows<<<nblk(a.n), kB, 0, s>>>(a.n, k[cur], vmask);
      GMPT_LAUNCH_CHECK();
    }
    e = radix_sort_u64(k[cur], v[cur], k[cur ^ 1], v[cur ^ 1], a.n, lo, 64, &which, temp, s);
    if (e != hipSuccess) return e;
    cur ^= which;
  } else {
    // LSD over (words..., length): length first, then the words from last to first
    int nwords = full;
    for (int w = -1; w < nwords; w++) {
      int word = w < 0 ? -1 : nwords - 1 - w;
      k_gather_word<<<nblk(a.n), kB, 0, s>>>(a, v[cur], k[cur], word);
      GMPT_LAUNCH_CHECK();
      e = radix_sort_u64(k[cur], v[cur], k[cur ^ 1], v[cur ^ 1], a.n, 0, word < 0 ? 16 : 64, &which, temp, s);
      if (e != hipSuccess) return e;
      cur ^= which;
    }
  }
  if (trie_bits > 0) {
    // stable sort of the current permutation by trie id, then regather the
    // windows in the new order
    uint32_t *t[2] = {a.tk0, a.tk1};
    k_gather_trie<<<nblk(a.n), kB, 0, s>>>(a, v[cur], t[0]);
    GMPT_LAUNCH_CHECK();
    e = radix_sort_u32(t[0], v[cur], t[1], v[cur ^ 1], a.n, 0, (int)((trie_bits + 7) & ~7u), &which, temp, s);
    if (e != hipSuccess) return e;
    if (which) cur ^= 1;
    if (!full) {
      k_gather_word<<<nblk(a.n), kB, 0, s>>>(a, v[cur], k[cur], 0);
      GMPT_LAUNCH_CHECK();
    }
  }
  *sk = k[cur];
  *sv = v[cur];
  if (cur_io) *cur_io = cur;
  // the trie pass regathered raw windows
  if (sk_mask) *sk_mask = (!full && win_all_bits && vmask && __builtin_popcountll(vmask) < 64 && trie_bits == 0)
                              ? vmask : 0ull;
  return hipSuccess;
}

hipError_t nu_adjacent(const NuArgs &a, const uint64_t *sk, const uint32_t *sv, int full, hipStream_t s,
                       uint64_t sk_mask) {
  k_adjacent<<<nblk(a.n), kB, 0, s>>>(a, sk, sv, full, sk_mask);
  return hipGetLastError();
}

namespace {
__global__ void k_reset_dup(uint32_t *scal) { scal[SC_DUP] = 0; }
}  // namespace

hipError_t nu_reset_dup(const NuArgs &a, hipStream_t s) {
  k_reset_dup<<<1, 1, 0, s>>>(a.scal);
  return hipGetLastError();
}

hipError_t nu_unique(const NuArgs &a, const uint64_t *sk, const uint32_t *sv, void *temp, hipStream_t s) {
  // run[] holds the run starts per block of k_adjacent; with many blocks
  // they are scanned here, else each block of k_unique adds up those before it
  const uint32_t nb = nblk(a.n);
  const bool pre = nb > kSelfSumBlocks;
  if (pre) {
    hipError_t e = scan_exclusive_u32(a.run, a.run, nb, nullptr, temp, s);
    if (e != hipSuccess) return e;
  }
  k_unique<<<nb, kB, 0, s>>>(a, sk, sv, a.run, pre ? 1 : 0);
  return hipGetLastError();
}

static hipError_t nu_groups(const NuArgs &a, void *temp, hipStream_t s) {
  k_group_flags<<<nblk(a.n), kB, 0, s>>>(a);
  GMPT_LAUNCH_CHECK();
  hipError_t e = scan_exclusive_u32(a.ecount, a.grp, a.n, nullptr, temp, s);
  if (e != hipSuccess) return e;
  k_group_fix<<<nblk(a.n), kB, 0, s>>>(a);
  GMPT_LAUNCH_CHECK();
  k_group_info<<<nblk(a.n), kB, 0, s>>>(a);
  GMPT_LAUNCH_CHECK();
  k_group_split<<<nblk(a.n), kB, 0, s>>>(a);
  return hipGetLastError();
}

hipError_t nu_descend_and_group(const NuArgs &a, void *temp, hipStream_t s, bool empty) {
  if (empty) {  // bulk build into empty tries: groups and elements in one pass
    const int multi = a.tid != nullptr;
    if (multi) {
      k_empty_flags<<<nblk(a.n), kB, 0, s>>>(a);
      GMPT_LAUNCH_CHECK();
      hipError_t e = scan_exclusive_u32(a.ecount, a.grp, a.n, nullptr, temp, s);
      if (e != hipSuccess) return e;
    }
    k_empty_groups<<<nblk(a.n), kB, 0, s>>>(a, multi);
    return hipGetLastError();
  }
  k_descend<<<nblk(a.n), kB, 0, s>>>(a);
  GMPT_LAUNCH_CHECK();
  hipError_t e = nu_groups(a, temp, s);
  if (e != hipSuccess) return e;
  k_group_merge<<<nblk(a.n), kB, 0, s>>>(a);
  return hipGetLastError();
}

hipError_t nu_regroup(NuArgs &a, uint32_t m, uint32_t nruns, void *temp, hipStream_t s) {
  k_gkey<<<nblk(m), kB, 0, s>>>(a, m);
  GMPT_LAUNCH_CHECK();
  int bits = 8;
  while (bits < 32 && (1ull << bits) < nruns) bits += 8;
  int which = 0;
  hipError_t e = radix_sort_u32(a.gkey0, a.perm0, a.gkey1, a.perm1, m, 0, bits, &which, temp, s);
  if (e != hipSuccess) return e;
  k_permute<<<nblk(m), kB, 0, s>>>(a, which ? a.perm1 : a.perm0, m);
  GMPT_LAUNCH_CHECK();
  uint32_t *t;
  t = a.u_idx; a.u_idx = a.u_idx2; a.u_idx2 = t;
  uint64_t *t64 = a.u_kref; a.u_kref = a.u_kref2; a.u_kref2 = t64;
  t64 = a.u_win; a.u_win = a.u_win2; a.u_win2 = t64;
  t = a.u_trie; a.u_trie = a.u_trie2; a.u_trie2 = t;
  uint8_t *l = a.u_lcp; a.u_lcp = a.u_lcp2; a.u_lcp2 = l;
  Stop *st = a.stop; a.stop = a.stop2; a.stop2 = st;
  return nu_groups(a, temp, s);
}

hipError_t nu_emit(const NuArgs &a, void *temp, hipStream_t s, bool empty) {
  if (empty) return hipSuccess;  // k_empty_groups emitted the elements
  k_emit_count<<<nblk(a.n), kB, 0, s>>>(a);
  GMPT_LAUNCH_CHECK();
  hipError_t e = scan_exclusive_u32(a.ecount, a.ecount, a.n, nullptr, temp, s);
  if (e != hipSuccess) return e;
  k_emit_write<<<nblk(a.n), kB, 0, s>>>(a);
  return hipGetLastError();
}

hipError_t nu_build(const NuArgs &a, void *temp, hipStream_t s, bool empty) {
  // elements: the unique keys, plus one virtual element per split short node
  // (none when the tries were empty)
  uint32_t ne_max = empty ? a.n + 1 : 2 * a.n + 2;
  if (empty) {  // k_empty_groups wrote levels 0 .. 2
    if (a.st_levels > 3) {
      k_mt_top<<<1, kB, 0, s>>>(a, 3);
      GMPT_LAUNCH_CHECK();
    }
  } else {
    size_t cnt = ne_max;  // bound on the entries of each level
    for (int k = 1; k < a.st_levels; k++) {
      cnt = (cnt + 15) / 16;
      k_mt_level<<<nblk(cnt), kB, 0, s>>>(a, k);
      GMPT_LAUNCH_CHECK();
    }
  }
  if (!empty) {  // empty tries hold no key to update
    k_updates<<<nblk(a.n), kB, 0, s>>>(a);
    GMPT_LAUNCH_CHECK();
  }
  k_split_reps<<<nblk(ne_max), kB, 0, s>>>(a, ne_max);
  GMPT_LAUNCH_CHECK();
  // the per-block representative counts: scanned here when there are many
  // blocks, else each block of k_build_links adds up those before it
  const bool pre = nblk(ne_max) > kSelfSumBlocks;
  if (pre) {
    hipError_t e = scan_exclusive_u32(a.bcnt, a.bcnt, nblk(ne_max), a.scal + SC_NREP_TOT, temp, s);
    if (e != hipSuccess) return e;
  }
  k_build_links<<<nblk(ne_max), kB, 0, s>>>(a, pre ? 1 : 0);
  GMPT_LAUNCH_CHECK();
  k_link_cross<<<nblk(ne_max), kB, 0, s>>>(a);
  return hipGetLastError();
}

}  // namespace gmpt
