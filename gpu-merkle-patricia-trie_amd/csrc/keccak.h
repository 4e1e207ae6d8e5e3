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

// 64-bit rotate as two 32-bit funnel shifts (v_alignbit_b32); r is a
// compile-time constant after unrolling.
__device__ __forceinline__ uint64_t rotl64(uint64_t v, int r) {
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  if (r == 0) return v;
  if (r >= 32) {
    uint32_t t = lo;
    lo = hi;
    hi = t;
    r -= 32;
    if (r == 0) return ((uint64_t)hi << 32) | lo;
  }
  uint32_t nlo = __builtin_amdgcn_alignbit(lo, hi, 32 - r);
  uint32_t nhi = __builtin_amdgcn_alignbit(hi, lo, 32 - r);
  return ((uint64_t)nhi << 32) | nlo;
}

// gfx950 three-input boolean op: result bit = imm bit (4 a + 2 b + c).
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}
// chi: a ^ (~b & c)
__device__ __forceinline__ uint32_t chi3(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0xD2);
}

// rotate-left of the 64-bit lane (lo, hi) by a compile-time r
__device__ __forceinline__ void rotl2(uint32_t lo, uint32_t hi, int r, uint32_t &olo, uint32_t &ohi) {
  if (r >= 32) {
    uint32_t t = lo;
    lo = hi;
    hi = t;
    r -= 32;
  }
  if (r == 0) {
    olo = lo;
    ohi = hi;
    return;
  }
  olo = __builtin_amdgcn_alignbit(lo, hi, 32 - r);
  ohi = __builtin_amdgcn_alignbit(hi, lo, 32 - r);
}

// Keccak-f[1600] on a state held in registers; lane A[x,y] is s[x + 5y].
// Per round and 32-bit half: theta's column parities with two xor3 each, the
// theta update fused into one xor3 per lane (A ^ C[x-1] ^ rot(C[x+1], 1)),
// rho as alignbit pairs, chi as one bitop3 per lane: 180 VALU ops per round.
__device__ __forceinline__ void keccak_f1600(uint64_t s[25]) {
  uint32_t L[25], H[25];
#pragma unroll
  for (int i = 0; i < 25; i++) {
    L[i] = (uint32_t)s[i];
    H[i] = (uint32_t)(s[i] >> 32);
  }
#pragma unroll 1
  for (int ir = 0; ir < 24; ir++) {
    uint32_t CL[5], CH[5], RL[5], RH[5], BL[25], BH[25];
#pragma unroll
    for (int x = 0; x < 5; x++) {
      CL[x] = xor3(xor3(L[x], L[x + 5], L[x + 10]), L[x + 15], L[x + 20]);
      CH[x] = xor3(xor3(H[x], H[x + 5], H[x + 10]), H[x + 15], H[x + 20]);
    }
#pragma unroll
    for (int x = 0; x < 5; x++) rotl2(CL[(x + 1) % 5], CH[(x + 1) % 5], 1, RL[x], RH[x]);
#pragma unroll
    for (int x = 0; x < 5; x++)
#pragma unroll
      for (int y = 0; y < 5; y++) {
        int i = x + 5 * y;
        uint32_t al = xor3(L[i], CL[(x + 4) % 5], RL[x]);
        uint32_t ah = xor3(H[i], CH[(x + 4) % 5], RH[x]);
        int j = y + 5 * ((2 * x + 3 * y) % 5);
        rotl2(al, ah, kKeccak.rho[i], BL[j], BH[j]);
      }
#pragma unroll
    for (int y = 0; y < 5; y++)
#pragma unroll
      for (int x = 0; x < 5; x++) {
        int i = x + 5 * y, i1 = (x + 1) % 5 + 5 * y, i2 = (x + 2) % 5 + 5 * y;
        L[i] = chi3(BL[i], BL[i1], BL[i2]);
        H[i] = chi3(BH[i], BH[i1], BH[i2]);
      }
    uint64_t rc = kKeccak.rc[ir];
    L[0] ^= (uint32_t)rc;
    H[0] ^= (uint32_t)(rc >> 32);
  }
#pragma unroll
  for (int i = 0; i < 25; i++) s[i] = ((uint64_t)H[i] << 32) | L[i];
}

}  // namespace gmpt
