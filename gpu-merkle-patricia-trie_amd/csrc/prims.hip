// prims.hip -- device-wide exclusive scan and stable LSD radix sort.
//
// Radix sort: reduce-then-scan with 8-bit digits.  Each block owns a tile of
// 1024 items (256 threads x 4 items, striped so loads are coalesced).  The
// scatter kernel ranks the tile stably in LDS (per-wave digit match via 8
// ballots), sorts it locally, then writes runs of equal digits contiguously.
#include "prims.h"

#include <algorithm>
#include "common.h"

namespace gmpt {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;
constexpr size_t kScanDirectTiles = 2048;  // up to this many tiles: two launches per scan

// With range (a device pair lo, hi), entries outside [lo, hi) count as zero
// and are neither read nor written (tiles outside the range touch no memory).
struct ScanRange {
  size_t lo, hi;
  __device__ __forceinline__ ScanRange(const uint32_t *range, size_t n)
      : lo(range ? range[0] : 0), hi(range ? (range[1] < n ? range[1] : n) : n) {}
  __device__ __forceinline__ bool in(size_t j) const { return j >= lo && j < hi; }
};

template <typename T>
__global__ __launch_bounds__(256) void k_scan_reduce(const T *__restrict__ in, size_t n, T *__restrict__ sums,
                                                     const uint32_t *skip, const uint32_t *range) {
  if (skip && *skip) return;  // uniform over the grid
  __shared__ T sh[4];
  const ScanRange rg(range, n);
  size_t base = (size_t)blockIdx.x * kScanTile;
  T acc = 0;
  if (base < rg.hi && base + kScanTile > rg.lo) {
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
      size_t j = base + (size_t)i * kScanBlock + threadIdx.x;
      if (rg.in(j)) acc += in[j];
    }
  }
  T tot;
  block_exclusive_scan_256<T>(acc, sh, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Scans one tile; its offset is offsets[blockIdx.x] (already scanned), or,
// with tile_sums, the sum of tile_sums[0 .. blockIdx.x) computed here (used
// when there are few tiles: saves the launch that scans the tile sums).
template <typename T>
__global__ __launch_bounds__(256) void k_scan_down(const T *__restrict__ in, T *__restrict__ out, size_t n,
                                                   const T *__restrict__ offsets, const T *__restrict__ tile_sums,
                                                   T *__restrict__ total, const uint32_t *skip,
                                                   const uint32_t *range) {
  if (skip && *skip) return;  // uniform over the grid
  __shared__ T tile[kScanTile];
  __shared__ T sh[4];
  const ScanRange rg(range, n);
  size_t base = (size_t)blockIdx.x * kScanTile;
  const bool touch = base < rg.hi && base + kScanTile > rg.lo;  // uniform over the block
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    size_t j = base + (size_t)i * kScanBlock + threadIdx.x;
    tile[i * kScanBlock + threadIdx.x] = touch && rg.in(j) ? in[j] : T(0);
  }
  T off = 0;
  if (tile_sums) {
    T part = 0;
    for (uint32_t j = threadIdx.x; j < blockIdx.x; j += kScanBlock) part += tile_sums[j];
    block_exclusive_scan_256<T>(part, sh, &off);
  } else if (offsets) {
    off = offsets[blockIdx.x];
  }
  __syncthreads();
  T v[kScanItems];
  T acc = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    v[i] = tile[threadIdx.x * kScanItems + i];
    acc += v[i];
  }
  T tot;
  T pre = block_exclusive_scan_256<T>(acc, sh, &tot);
  pre += off;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    tile[threadIdx.x * kScanItems + i] = pre;
    pre += v[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    size_t j = base + (size_t)i * kScanBlock + threadIdx.x;
    if (touch && rg.in(j)) out[j] = tile[i * kScanBlock + threadIdx.x];
  }
  if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total = off + tot;
}

// A scan of at most kScanSmallTiles tiles in one workgroup: each tile as
// k_scan_down scans it, offset by the tiles before (one launch in place of
// the reduce + scan pair).
constexpr size_t kScanSmallTiles = 4;
template <typename T>
__global__ __launch_bounds__(256) void k_scan_small(const T *__restrict__ in, T *__restrict__ out, size_t n,
                                                    T *__restrict__ total, const uint32_t *skip,
                                                    const uint32_t *range, uint32_t ntiles) {
  if (skip && *skip) return;  // uniform over the grid
  __shared__ T tile[kScanTile];
  __shared__ T sh[4];
  const ScanRange rg(range, n);
  T off = 0;
  for (uint32_t b = 0; b < ntiles; b++) {
    const size_t base = (size_t)b * kScanTile;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
      const size_t j = base + (size_t)i * kScanBlock + threadIdx.x;
      tile[i * kScanBlock + threadIdx.x] = rg.in(j) ? in[j] : T(0);
    }
    __syncthreads();
    T v[kScanItems];
    T acc = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
      v[i] = tile[threadIdx.x * kScanItems + i];
      acc += v[i];
    }
    T tot;
    T pre = block_exclusive_scan_256<T>(acc, sh, &tot) + off;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
      tile[threadIdx.x * kScanItems + i] = pre;
      pre += v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
      const size_t j = base + (size_t)i * kScanBlock + threadIdx.x;
      if (rg.in(j)) out[j] = tile[i * kScanBlock + threadIdx.x];
    }
    __syncthreads();  // the tile buffer is refilled next
    off += tot;
  }
  if (total && threadIdx.x == 0) *total = off;
}

template <typename T>
size_t scan_temp_elems(size_t n) {
  size_t e = 0;
  while (n > (size_t)kScanTile) {
    n = (n + kScanTile - 1) / kScanTile;
    e += n + 1;
  }
  return e + 1;
}

template <typename T>
hipError_t scan_impl(const T *in, T *out, size_t n, T *total, T *temp, hipStream_t s,
                     const uint32_t *skip = nullptr, const uint32_t *range = nullptr) {
  if (n == 0) {
    if (total) return hipMemsetAsync(total, 0, sizeof(T), s);
    return hipSuccess;
  }
  size_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb <= kScanSmallTiles) {  // one workgroup, tile after tile
    k_scan_small<T><<<1, kScanBlock, 0, s>>>(in, out, n, total, skip, range, (uint32_t)nb);
    return hipGetLastError();
  }
  T *sums = temp;
  T *rest = temp + nb + 1;
  k_scan_reduce<T><<<(unsigned)nb, kScanBlock, 0, s>>>(in, n, sums, skip, range);
  GMPT_LAUNCH_CHECK();
  if (nb <= kScanDirectTiles) {  // each tile adds up the sums before it
    k_scan_down<T><<<(unsigned)nb, kScanBlock, 0, s>>>(in, out, n, nullptr, sums, total, skip, range);
    return hipGetLastError();
  }
  hipError_t e = scan_impl<T>(sums, sums, nb, nullptr, rest, s, skip);
  if (e != hipSuccess) return e;
  k_scan_down<T><<<(unsigned)nb, kScanBlock, 0, s>>>(in, out, n, sums, nullptr, total, skip, range);
  return hipGetLastError();
}

// ------------------------------------------------------------ radix sort --

constexpr int kRsBlock = 256;
constexpr int kRsItems = 4;
constexpr int kRsTile = kRsBlock * kRsItems;
constexpr int kRsWaves = kRsBlock / 64;

template <typename K>
__global__ __launch_bounds__(256) void k_rs_hist(const K *__restrict__ keys, size_t n, int shift,
                                                 uint32_t *__restrict__ hist, uint32_t nblocks) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  size_t base = (size_t)blockIdx.x * kRsTile;
#pragma unroll
  for (int r = 0; r < kRsItems; r++) {
    size_t j = base + (size_t)r * kRsBlock + threadIdx.x;
    if (j < n) atomicAdd(&h[(uint32_t)(keys[j] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

template <typename K>
__global__ __launch_bounds__(256) void k_rs_scatter(const K *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                    K *__restrict__ kout, uint32_t *__restrict__ vout, size_t n,
                                                    int shift, const uint32_t *__restrict__ goff,
                                                    uint32_t nblocks) {
  __shared__ uint16_t cnt[kRsItems][kRsWaves][256];
  __shared__ K sk[kRsTile];
  __shared__ uint32_t sv[kRsTile];
  __shared__ uint32_t tstart[256];
  __shared__ uint32_t shs[4];

  const int t = threadIdx.x;
  const int lane = __lane_id();
  const int w = t >> 6;
  const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  size_t base = (size_t)blockIdx.x * kRsTile;

  for (int i = t; i < kRsItems * kRsWaves * 256; i += kRsBlock) (&cnt[0][0][0])[i] = 0;
  __syncthreads();

  K key[kRsItems];
  uint32_t val[kRsItems], dig[kRsItems], wrank[kRsItems];
  bool valid[kRsItems];
#pragma unroll
  for (int r = 0; r < kRsItems; r++) {
    size_t j = base + (size_t)r * kRsBlock + t;
    valid[r] = j < n;
    key[r] = valid[r] ? kin[j] : K(0);
    val[r] = valid[r] ? vin[j] : 0u;
    dig[r] = (uint32_t)(key[r] >> shift) & 255u;
    uint64_t m = __ballot(valid[r]);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      uint64_t bb = __ballot((dig[r] >> b) & 1u);
      m &= ((dig[r] >> b) & 1u) ? bb : ~bb;
    }
    wrank[r] = (uint32_t)__popcll(m & lt_mask);
    // the lowest lane of each digit group records the group's size
    if (valid[r] && (m & lt_mask) == 0) cnt[r][w][dig[r]] = (uint16_t)__popcll(m);
  }
  __syncthreads();
  // per digit: exclusive prefix over (round, wave) in stable order
  {
    const int d = t;
    uint32_t run = 0;
#pragma unroll
    for (int r = 0; r < kRsItems; r++)
#pragma unroll
      for (int ww = 0; ww < kRsWaves; ww++) {
        uint32_t c = cnt[r][ww][d];
        cnt[r][ww][d] = (uint16_t)run;
        run += c;
      }
    uint32_t tot;
    uint32_t st = block_exclusive_scan_256<uint32_t>(run, shs, &tot);
    tstart[d] = st;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRsItems; r++) {
    if (valid[r]) {
      uint32_t pos = tstart[dig[r]] + cnt[r][w][dig[r]] + wrank[r];
      sk[pos] = key[r];
      sv[pos] = val[r];
    }
  }
  __syncthreads();
  size_t valid_n = n - base < (size_t)kRsTile ? n - base : (size_t)kRsTile;
#pragma unroll
  for (int r = 0; r < kRsItems; r++) {
    uint32_t i = r * kRsBlock + t;
    if (i < valid_n) {
      K k = sk[i];
      uint32_t d = (uint32_t)(k >> shift) & 255u;
      size_t dst = (size_t)goff[(size_t)d * nblocks + blockIdx.x] + (i - tstart[d]);
      kout[dst] = k;
      vout[dst] = sv[i];
    }
  }
}

template <typename K>
hipError_t radix_impl(K *k0, uint32_t *v0, K *k1, uint32_t *v1, size_t n, int begin_bit, int end_bit, int *which,
                      void *temp, hipStream_t s) {
  *which = 0;
  if (n <= 1) return hipSuccess;
  uint32_t nb = (uint32_t)((n + kRsTile - 1) / kRsTile);
  uint32_t *hist = (uint32_t *)temp;
  uint32_t *scan_tmp = hist + (size_t)256 * nb + 16;
  K *kin = k0, *kout = k1;
  uint32_t *vin = v0, *vout = v1;
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    k_rs_hist<K><<<nb, kRsBlock, 0, s>>>(kin, n, shift, hist, nb);
    GMPT_LAUNCH_CHECK();
    hipError_t e = scan_impl<uint32_t>(hist, hist, (size_t)256 * nb, nullptr, scan_tmp, s);
    if (e != hipSuccess) return e;
    k_rs_scatter<K><<<nb, kRsBlock, 0, s>>>(kin, vin, kout, vout, n, shift, hist, nb);
    GMPT_LAUNCH_CHECK();
    K *tk = kin;
    kin = kout;
    kout = tk;
    uint32_t *tv = vin;
    vin = vout;
    vout = tv;
    *which ^= 1;
  }
  return hipSuccess;
}

// One bucket of the top-8-bit pass per workgroup: bucket d spans
// [goff[d * nblocks], goff[(d + 1) * nblocks]) of the pass's output (the
// exclusive scan of the digit-major per-tile counts), or the whole array
// without goff.  The bucket is counting-sorted in LDS by the next 12 key bits
// (4,096 sub-buckets: about one entry each for a bucket of uniform keys),
// then each sub-bucket's run is insertion-sorted by (key, value), which also
// fixes the order the LDS atomics placed equal sub-buckets in.  A bucket
// with a run longer than kSubMaxRun is left as it came and sets bit 1 of
// *overflow.
constexpr int kBkThreads = 1024;
constexpr uint32_t kBkPer = (uint32_t)(kBucketCap / kBkThreads);  // entries per thread
constexpr uint32_t kSubs = 4096;
constexpr uint32_t kSubPer = kSubs / kBkThreads;                    // sub-buckets per thread
constexpr uint32_t kSubMaxRun = 64;
static_assert(kBucketCap % kBkThreads == 0 && kSubs % kBkThreads == 0, "bucket tiling");

__global__ __launch_bounds__(1024) void k_bucket_sort(uint64_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                      uint32_t n, const uint32_t *__restrict__ goff,
                                                      uint32_t nblocks, uint32_t *overflow) {
  __shared__ uint64_t sk[kBucketCap];
  __shared__ uint32_t sv[kBucketCap];
  __shared__ uint32_t start[kSubs];
  __shared__ uint32_t wsum[kBkThreads / 64];
  const uint32_t d = blockIdx.x;
  const uint32_t lo = goff ? goff[(size_t)d * nblocks] : 0u;
  const uint32_t hi = goff && d + 1 < 256 ? goff[(size_t)(d + 1) * nblocks] : n;
  const uint32_t c = hi - lo;  // uniform over the block
  if (c <= 1) return;
  if (c > kBucketCap) {
    if (threadIdx.x == 0) atomicOr(overflow, 2u);
    return;
  }
  // the sub-bucket bits: below the bucket's 8, or the top 12 of a single bucket
  const int sub_shift = goff ? 44 : 52;
  const uint32_t t = threadIdx.x;
#pragma unroll
  for (uint32_t j = 0; j < kSubPer; j++) start[t + j * kBkThreads] = 0;
  __syncthreads();
  uint64_t k[kBkPer];
  uint32_t v[kBkPer], sub[kBkPer], pos[kBkPer];
#pragma unroll
  for (uint32_t q = 0; q < kBkPer; q++) {
    const uint32_t e = t + q * kBkThreads;
    if (e < c) {
      k[q] = keys[lo + e];
      v[q] = vals[lo + e];
      sub[q] = (uint32_t)(k[q] >> sub_shift) & (kSubs - 1);
      pos[q] = atomicAdd(&start[sub[q]], 1u);
    }
  }
  __syncthreads();
  // exclusive scan of the counts; thread t owns sub-buckets kSubPer t ..
  uint32_t cnt[kSubPer], sum = 0;
#pragma unroll
  for (uint32_t j = 0; j < kSubPer; j++) {
    cnt[j] = start[kSubPer * t + j];
    sum += cnt[j];
  }
  const uint32_t incl = wave_inclusive_scan(sum);
  if (__lane_id() == 63) wsum[t >> 6] = incl;
  __syncthreads();
  uint32_t base = 0;
  for (uint32_t w = 0; w < (t >> 6); w++) base += wsum[w];
  uint32_t first[kSubPer];
  bool long_run = false;
  {
    uint32_t x = base + incl - sum;
#pragma unroll
    for (uint32_t j = 0; j < kSubPer; j++) {
      first[j] = x;
      start[kSubPer * t + j] = x;  // only this thread reads or writes these entries here
      x += cnt[j];
      long_run |= cnt[j] > kSubMaxRun;
    }
  }
  // a run too long to insertion-sort: the bucket stays as it came (in the
  // previous pass's stable order, which the caller's other sort relies on)
  if (__syncthreads_or(long_run)) {
    if (t == 0) atomicOr(overflow, 2u);
    return;
  }
#pragma unroll
  for (uint32_t q = 0; q < kBkPer; q++) {
    const uint32_t e = t + q * kBkThreads;
    if (e < c) {
      const uint32_t p = start[sub[q]] + pos[q];
      sk[p] = k[q];
      sv[p] = v[q];
    }
  }
  __syncthreads();
#pragma unroll
  for (uint32_t j = 0; j < kSubPer; j++) {
    if (cnt[j] < 2) continue;
    const uint32_t f = first[j];
    for (uint32_t q = 1; q < cnt[j]; q++) {
      const uint64_t kk = sk[f + q];
      const uint32_t vv = sv[f + q];
      uint32_t p = q;
      while (p > 0 && (sk[f + p - 1] > kk || (sk[f + p - 1] == kk && sv[f + p - 1] > vv))) {
        sk[f + p] = sk[f + p - 1];
        sv[f + p] = sv[f + p - 1];
        p--;
      }
      sk[f + p] = kk;
      sv[f + p] = vv;
    }
  }
  __syncthreads();
  for (uint32_t e = t; e < c; e += kBkThreads) {
    keys[lo + e] = sk[e];
    vals[lo + e] = sv[e];
  }
}

}  // namespace

hipError_t bucket_sort_u64(uint64_t *k0, uint32_t *v0, uint64_t *k1, uint32_t *v1, size_t n, int *which, void *temp,
                           hipStream_t s, uint32_t *overflow) {
  *which = 0;
  if (n <= 1) return hipSuccess;
  if (n > kBucketSortMaxN) return hipErrorInvalidValue;
  if (n <= kBucketCap) {
    k_bucket_sort<<<1, kBkThreads, 0, s>>>(k0, v0, (uint32_t)n, nullptr, 0, overflow);
    return hipGetLastError();
  }
  const uint32_t nb = (uint32_t)((n + kRsTile - 1) / kRsTile);
  uint32_t *hist = (uint32_t *)temp;
  uint32_t *scan_tmp = hist + (size_t)256 * nb + 16;
  k_rs_hist<uint64_t><<<nb, kRsBlock, 0, s>>>(k0, n, 56, hist, nb);
  GMPT_LAUNCH_CHECK();
  hipError_t e = scan_impl<uint32_t>(hist, hist, (size_t)256 * nb, nullptr, scan_tmp, s);
  if (e != hipSuccess) return e;
  k_rs_scatter<uint64_t><<<nb, kRsBlock, 0, s>>>(k0, v0, k1, v1, n, 56, hist, nb);
  GMPT_LAUNCH_CHECK();
  k_bucket_sort<<<256, kBkThreads, 0, s>>>(k1, v1, (uint32_t)n, hist, nb, overflow);
  GMPT_LAUNCH_CHECK();
  *which = 1;
  return hipSuccess;
}

size_t scan_temp_bytes(size_t n) { return scan_temp_elems<uint64_t>(n) * sizeof(uint64_t) + 64; }

size_t radix_temp_bytes(size_t n) {
  size_t nb = (n + kRsTile - 1) / kRsTile + 1;
  return (256 * nb + 16) * sizeof(uint32_t) + scan_temp_bytes(256 * nb) + 256;
}

hipError_t scan_exclusive_u32(const uint32_t *in, uint32_t *out, size_t n, uint32_t *total, void *temp,
                              hipStream_t s) {
  return scan_impl<uint32_t>(in, out, n, total, (uint32_t *)temp, s);
}

hipError_t scan_exclusive_u64(const uint64_t *in, uint64_t *out, size_t n, uint64_t *total, void *temp,
                              hipStream_t s, const uint32_t *skip, const uint32_t *range) {
  return scan_impl<uint64_t>(in, out, n, total, (uint64_t *)temp, s, skip, range);
}

hipError_t radix_sort_u64(uint64_t *k0, uint32_t *v0, uint64_t *k1, uint32_t *v1, size_t n, int begin_bit,
                          int end_bit, int *which, void *temp, hipStream_t s) {
  return radix_impl<uint64_t>(k0, v0, k1, v1, n, begin_bit, end_bit, which, temp, s);
}

hipError_t radix_sort_u32(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, size_t n, int begin_bit,
                          int end_bit, int *which, void *temp, hipStream_t s) {
  return radix_impl<uint32_t>(k0, v0, k1, v1, n, begin_bit, end_bit, which, temp, s);
}

namespace {
__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace

hipError_t device_copy16(void *dst, const void *src, size_t bytes, hipStream_t s) {
  size_t n16 = bytes / 16;
  if (n16 == 0) return hipSuccess;
  size_t blocks = std::min<size_t>((n16 + 255) / 256, 8192);
  k_copy16<<<(unsigned)blocks, 256, 0, s>>>((const uint4 *)src, (uint4 *)dst, n16);
  return hipGetLastError();
}

}  // namespace gmpt
