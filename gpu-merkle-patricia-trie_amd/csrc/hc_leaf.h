// hc_leaf.h -- a leaf short node's hashing bin and metadata, shared by the
// PhaseHC classification (which derives them from the nodes) and the PhaseNU
// build (which records them for the leaves it creates, so the classification
// need not chase those nodes).
#pragma once
#include "common.h"
#include "phasehc.h"

namespace gmpt {

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

// Keccak-256 blocks (rate 136 bytes) of a message of len bytes plus padding.
__device__ __forceinline__ uint32_t keccak_blocks(uint32_t len) { return len / 136 + 1; }

// Keccak blocks of a leaf [compact key of cnt nibbles, value of vsz bytes],
// bounded from above: a 1-byte value counts with its string header.
__device__ __forceinline__ uint32_t leaf_blocks(uint32_t cnt, uint32_t vsz) {
  const uint32_t pay = rlp_item_len(cnt / 2 + 1, 0) + rlp_item_len(vsz, 0x80);
  return keccak_blocks(rlp_hdr_len(pay) + pay);
}

// Leaf metadata for the leaf-level kernel:
//   m0 = byte offset of the key's first packed byte in the key arena (40 bits)
//        | key_start & 1 << 40 | key nibbles without terminator << 41 (8 bits)
//        | term << 49 | value size << 50 (14 bits, saturated)
//   m1 = device address of the value bytes
__device__ __forceinline__ void leaf_meta_words(uint64_t key_off, uint32_t key_start, uint32_t cnt, uint32_t term,
                                                uint32_t vsz, uint64_t va, uint64_t *m) {
  m[0] = (key_off + key_start / 2) | ((uint64_t)(key_start & 1) << 40) | ((uint64_t)cnt << 41) |
         ((uint64_t)term << 49) | ((uint64_t)min(vsz, 16383u) << 50);
  m[1] = va;
}

constexpr uint32_t kLeafPackedNone = 0xffffffffu;  // no recorded bin: derive it from the node
constexpr uint32_t kPackedNbMax = (1u << 19) - 1;   // blocks field of a packed bin, saturated
constexpr uint32_t kLeafFastNibbles = 64;           // key nibbles a leaf head holds in registers
constexpr uint32_t kLeafDirectMaxValue = 16382;     // value bytes the metadata holds (14 bits, saturated above)

// Packed bin (bin | blocks << 12) of a leaf: class by blocks, level 0 when
// the leaf-level word builder takes it (key remainder <= 64 nibbles, value
// <= 16,382 bytes), else level 1 (the generic leaf encoder).
__device__ __forceinline__ uint32_t leaf_packed(uint32_t cnt, uint32_t vsz) {
  const uint32_t nb = leaf_blocks(cnt, vsz);
  const uint32_t cls = nb > kHcClasses ? kHcClasses - 1 : nb - 1;
  const uint32_t level = cnt > kLeafFastNibbles || vsz > kLeafDirectMaxValue ? 1u : 0u;
  return (level * kHcClasses + cls) | (min(nb, kPackedNbMax) << 12);
}

// A new branch's record (NuArgs::bdesc): depth | children << 16, and
// kBranchRecChase when a child sits in slot 16 (its value's size counts).
constexpr uint32_t kBranchRecChild = 1u << 16;
constexpr uint32_t kBranchRecChase = 1u << 31;

// Packed bin of a branch from its record, as the classification derives it
// from the node (every child counted as a 32-byte digest, slot 16 empty), or
// kLeafPackedNone when the node must be read.
__device__ __forceinline__ uint32_t branch_packed(uint32_t rec) {
  if (rec & kBranchRecChase) return kLeafPackedNone;
  const uint32_t depth = rec & 0xffffu, k = (rec >> 16) & 0x7fffu;
  const uint32_t pay = 17 + 32 * k;  // 16 - k empty slots, k 33-byte children, empty slot 16
  const uint32_t nb = keccak_blocks(rlp_hdr_len(pay) + pay);
  const uint32_t cls = nb > kHcClasses ? kHcClasses - 1 : nb - 1;
  const uint32_t level = 1 + (kHcMaxDepth - min(depth, kHcMaxDepth));
  return (level * kHcClasses + cls) | (min(nb, kPackedNbMax) << 12);
}

}  // namespace gmpt
