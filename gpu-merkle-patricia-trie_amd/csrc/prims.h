// prims.h -- device-wide scan and LSD radix sort (hand-written, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gmpt {

// Bytes of temporary storage needed by the primitives below for n elements.
size_t scan_temp_bytes(size_t n);
size_t radix_temp_bytes(size_t n);

// out[i] = sum(in[0..i)) ; total (if not null) receives the sum of all n
// (device pointer).  `in` and `out` may alias.
hipError_t scan_exclusive_u32(const uint32_t *in, uint32_t *out, size_t n, uint32_t *total, void *temp,
                              hipStream_t s);
// With skip (a device word), the scan does nothing when *skip != 0.  With
// range (device words lo, hi), only entries [lo, hi) are summed, read and
// written; out[i] for i in [lo, hi) is the sum of in[lo .. i).
hipError_t scan_exclusive_u64(const uint64_t *in, uint64_t *out, size_t n, uint64_t *total, void *temp,
                              hipStream_t s, const uint32_t *skip = nullptr, const uint32_t *range = nullptr);

// Stable LSD sort of (keys, vals) by key bits [begin_bit, end_bit), 8 bits per
// pass.  Buffers ping-pong between (k0,v0) and (k1,v1); returns in *which the
// buffer index (0 or 1) holding the result.
hipError_t radix_sort_u64(uint64_t *k0, uint32_t *v0, uint64_t *k1, uint32_t *v1, size_t n, int begin_bit,
                          int end_bit, int *which, void *temp, hipStream_t s);
hipError_t radix_sort_u32(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, size_t n, int begin_bit,
                          int end_bit, int *which, void *temp, hipStream_t s);

// Sort of (keys, vals) by (key, val) for n <= kBucketSortMaxN: one stable
// pass on the key's top 8 bits, then each of the 256 buckets sorted in LDS by
// one workgroup (a counting sort on the next 12 bits, then an insertion sort
// of each run of equal 20-bit prefixes by (key, val), so equal keys end in
// value order).  A bucket above kBucketCap entries, or with a run above 64,
// is left in the first pass's stable order and sets bit 1 of *overflow: the
// caller then sorts another way, stably (keys of uniform spread fill buckets
// of ~n/256 and runs of ~1).
// n <= kBucketCap skips the first pass.  Buffers as radix_sort_u64.
constexpr size_t kBucketCap = 8192;
constexpr size_t kBucketSortMaxN = 1600000;  // 6,250 per bucket on average
hipError_t bucket_sort_u64(uint64_t *k0, uint32_t *v0, uint64_t *k1, uint32_t *v1, size_t n, int *which, void *temp,
                           hipStream_t s, uint32_t *overflow);

// ---------------------------------------------------------------- device --

// Inclusive scan of one value per lane over a wave64.
template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
  const int lane = __lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    T o = __shfl_up(v, off, 64);
    if (lane >= off) v += o;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread (blockDim = 256).
// `sh` must hold 4 elements of T.  Returns the exclusive prefix; *total gets
// the block sum.
template <typename T>
__device__ __forceinline__ T block_exclusive_scan_256(T v, T *sh, T *total) {
  const int lane = __lane_id();
  const int w = threadIdx.x >> 6;
  T inc = wave_inclusive_scan(v);
  if (lane == 63) sh[w] = inc;
  __syncthreads();
  T base = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    if (i < w) base += sh[i];
    tot += sh[i];
  }
  __syncthreads();
  if (total) *total = tot;
  return base + inc - v;
}

// Plain streaming copy of bytes / 16 16-byte words (a reference kernel of
// known traffic for calibrating memory counters).
hipError_t device_copy16(void *dst, const void *src, size_t bytes, hipStream_t s);

}  // namespace gmpt
