// keccak.h -- Keccak-f[1600] with one 25-lane state per GPU lane, held in VGPRs.
//
// The reference hashes with 25 of 32 warp lanes sharing one state in shared
// memory (mpt-with-compress/include/hash/batch_mode_hash.cuh:12-87) and has a
// single-thread variant (include/hash/gpu_hash.cuh:88-129).  On CDNA the whole
// 200-byte state fits in 50 VGPRs, so every lane of a wave64 runs its own
// message and no cross-lane traffic is needed; 64-bit rotations lower to pairs
// of v_alignbit_b32 and chi's andn+xor to v_bfi_b32-style selects.
//
// Round constants and rotation offsets are computed at compile time from
// FIPS 202 (Algorithm 5 for rc, Algorithm 2 for rho), not copied from tables.
#pragma once
#include <stdint.h>

namespace gmpt {

struct KeccakConsts {
  uint64_t rc[24];
  int rho[25];
  constexpr KeccakConsts() : rc{}, rho{} {
    // rc(t): LFSR x^8 + x^6 + x^5 + x^4 + 1, bit R[0] is the output.
    for (int ir = 0; ir < 24; ir++) {
      uint64_t v = 0;
      for (int j = 0; j <= 6; j++) {
        int t = j + 7 * ir;
        int bit = 1;
        if (t % 255 != 0) {
          unsigned r = 0x80;  // R[0] is bit 7
          for (int i = 1; i <= t % 255; i++) {
            unsigned r8 = r >> 1;            // R = 0 || R, truncated to 8 bits
            unsigned out = r & 1;            // old R[7] becomes R[8]
            if (out) r8 ^= 0x80 | 0x08 | 0x04 | 0x02;  // R[0], R[4], R[5], R[6] ^= R[8]
            r = r8;
          }
          bit = (r >> 7) & 1;
        }
        if (bit) v |= (uint64_t)1 << ((1 << j) - 1);
      }
      rc[ir] = v;
    }
    for (int i = 0; i < 25; i++) rho[i] = 0;
    int x = 1, y = 0;
    for (int t = 0; t < 24; t++) {
      rho[x + 5 * y] = ((t + 1) * (t + 2) / 2) % 64;
      int nx = y, ny = (2 * x + 3 * y) % 5;
      x = nx;
      y = ny;
    }
  }
};

constexpr KeccakConsts kKeccak{};
static_assert(kKeccak.rc[0] == 1ull, "rc[0]");
static_assert(kKeccak.rc[1] == 0x8082ull, "rc[1]");
static_assert(kKeccak.rho[1] == 1 && kKeccak.rho[2] == 62, "rho");

__device__ __forceinline__ uint64_t rotl64(uint64_t v, int r) {
  return r == 0 ? v : ((v << r) | (v >> (64 - r)));
}

// Keccak-f[1600] on a state held in registers; lane A[x,y] is s[x + 5y].
__device__ __forceinline__ void keccak_f1600(uint64_t s[25]) {
#pragma unroll 1
  for (int ir = 0; ir < 24; ir++) {
    uint64_t C[5], D[5], B[25];
#pragma unroll
    for (int x = 0; x < 5; x++) C[x] = s[x] ^ s[x + 5] ^ s[x + 10] ^ s[x + 15] ^ s[x + 20];
#pragma unroll
    for (int x = 0; x < 5; x++) D[x] = C[(x + 4) % 5] ^ rotl64(C[(x + 1) % 5], 1);
#pragma unroll
    for (int x = 0; x < 5; x++)
#pragma unroll
      for (int y = 0; y < 5; y++)
        B[y + 5 * ((2 * x + 3 * y) % 5)] = rotl64(s[x + 5 * y] ^ D[x], kKeccak.rho[x + 5 * y]);
#pragma unroll
    for (int y = 0; y < 5; y++)
#pragma unroll
      for (int x = 0; x < 5; x++)
        s[x + 5 * y] = B[x + 5 * y] ^ ((~B[(x + 1) % 5 + 5 * y]) & B[(x + 2) % 5 + 5 * y]);
    s[0] ^= kKeccak.rc[ir];
  }
}

}  // namespace gmpt
