// phasehc.h -- PhaseHC: bottom-up one-pass hashing (host launch interface).
#pragma once
#include "common.h"

namespace gmpt {

constexpr uint32_t kHcClasses = 8;     // Keccak block-count classes per level
constexpr uint32_t kHcMaxDepth = 255;  // levels: 0 = leaves, 1 + (255 - depth) otherwise
constexpr uint32_t kHcLevels = kHcMaxDepth + 2;
constexpr uint32_t kHcBins = kHcLevels * kHcClasses;

// Bins dirty nodes by (level, block class): hist[bin] = count, out = the list
// ordered by bin.
hipError_t hc_classify(const Pools &P, const ref_t *list, uint32_t n, uint32_t *hist, uint32_t *cursor,
                       ref_t *out, void *temp, hipStream_t s);
// Encodes and hashes list[0..n) (one level: no node depends on another).
hipError_t hc_hash_range(const Pools &P, const ref_t *list, uint32_t n, hipStream_t s);
// Copies every trie's root hash into its header.
hipError_t hc_roots(const Pools &P, uint32_t ntries, hipStream_t s);
// Node export: encoding lengths, then encodings at offs and 32-byte hash slots.
hipError_t hc_enc_len(const Pools &P, const ref_t *refs, uint32_t m, uint64_t *len, hipStream_t s);
hipError_t hc_enc_write(const Pools &P, const ref_t *refs, uint32_t m, const uint64_t *offs, uint8_t *encs,
                        uint8_t *hashes, hipStream_t s);
// Keccak-256 of n messages (test hook for the permutation and sponge).
hipError_t keccak_batch(const uint8_t *data, const uint64_t *off, const uint32_t *len, uint32_t n, uint8_t *out,
                        hipStream_t s);

}  // namespace gmpt
