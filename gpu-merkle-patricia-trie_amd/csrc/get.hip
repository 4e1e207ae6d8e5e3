// get.hip -- point lookups (gets_parallel) and path node export.
//
// Reference: gets_parallel / get (mpt-with-compress/include/mpt/
// gpu_mpt_kernels.cuh:2267-2340) returns the caller's host pointer recorded at
// insert and the value size; flush_dirty_nodes (gpu_mpt.cuh:1039-1122,
// kernels :2342-2588) exports (hash, encoding) of every node on the given
// keys' paths.
#include <string.h>

#include <vector>

#include "get.h"
#include "phasehc.h"
#include "prims.h"

namespace gmpt {

namespace {

constexpr int kB = 256;
__device__ __forceinline__ uint32_t grid_tid() { return blockIdx.x * blockDim.x + threadIdx.x; }
inline unsigned nblk(size_t n) { return (unsigned)((n + kB - 1) / kB); }

struct HexKey {
  const uint8_t *k;
  uint32_t len;
};

// The key's hex nibbles, or len 0 when its index pair is out of range.
__device__ __forceinline__ HexKey hex_key(const GetArgs &g, uint32_t i) {
  HexKey k;
  int32_t s = g.kidx[2 * i], e = g.kidx[2 * i + 1];
  bool ok = s >= 0 && e >= s && (uint64_t)e < g.kh_bytes;
  k.k = g.kh + (ok ? s : 0);
  k.len = ok ? (uint32_t)(e - s + 1) : 0u;
  return k;
}

__device__ __forceinline__ uint32_t short_nib(const Pools &P, const Short &S, uint32_t rel) {
  if (S.term && rel + 1 == S.key_len) return kNibbleT;
  return packed_nibble(P.keys + S.key_off, S.key_start + rel);
}

// Walks the key's path; calls visit(ref) for every branch / short node and
// returns the value node reached (REF_NULL when the key is absent).
template <typename F>
__device__ ref_t walk(const Pools &P, ref_t node, HexKey k, F visit) {
  uint32_t pos = 0;
  for (int guard = 0; guard < 2 * kMaxKeyNibbles + 4; guard++) {
    uint32_t t = ref_type(node);
    if (node == REF_NULL) return REF_NULL;
    if (t == T_VALUE) return pos == k.len ? node : REF_NULL;
    visit(node);
    if (t == T_BRANCH) {
      if (pos >= k.len) return REF_NULL;
      uint32_t nib = k.k[pos];
      if (nib > 16) return REF_NULL;
      node = P.branch[ref_idx(node)].child[nib];
      pos++;
    } else {
      const Short &S = P.shrt[ref_idx(node)];
      if (k.len - pos < S.key_len) return REF_NULL;
      for (uint32_t q = 0; q < S.key_len; q++)
        if (short_nib(P, S, q) != k.k[pos + q]) return REF_NULL;
      pos += S.key_len;
      node = S.child;
    }
  }
  return REF_NULL;
}

__global__ __launch_bounds__(256) void k_gets(GetArgs g) {
  uint32_t i = grid_tid();
  if (i >= g.n) return;
  HexKey k = hex_key(g, i);
  uint32_t t = g.tid ? (uint32_t)g.tid[i] : 0u;
  ref_t v = REF_NULL;
  if (t < g.ntries && k.len > 0) v = walk(g.P, g.P.tries[t].root, k, [](ref_t) {});
  if (v == REF_NULL) {
    g.hp[i] = 0;
    g.addr[i] = 0;
    g.size[i] = -1;
  } else {
    const Value &V = g.P.value[ref_idx(v)];
    g.hp[i] = V.hp;
    g.addr[i] = V.addr;
    g.size[i] = (int32_t)V.size;
  }
}

__global__ __launch_bounds__(256) void k_gather_values(const uint64_t *addr, const int32_t *size,
                                                       const uint64_t *off, uint32_t n, uint8_t *out) {
  uint32_t i = blockIdx.x;
  if (i >= n || size[i] <= 0) return;
  const uint8_t *src = (const uint8_t *)addr[i];
  for (int32_t q = threadIdx.x; q < size[i]; q += blockDim.x) out[off[i] + q] = src[q];
}

__global__ __launch_bounds__(256) void k_path_count(GetArgs g, uint32_t *cnt) {
  uint32_t i = grid_tid();
  if (i >= g.n) return;
  HexKey k = hex_key(g, i);
  uint32_t t = g.tid ? (uint32_t)g.tid[i] : 0u;
  uint32_t c = 0;
  if (t < g.ntries && k.len > 0) walk(g.P, g.P.tries[t].root, k, [&](ref_t) { c++; });
  cnt[i] = c;
}

__global__ __launch_bounds__(256) void k_path_write(GetArgs g, const uint32_t *off, ref_t *out, uint32_t *idx) {
  uint32_t i = grid_tid();
  if (i >= g.n) return;
  HexKey k = hex_key(g, i);
  uint32_t t = g.tid ? (uint32_t)g.tid[i] : 0u;
  uint32_t o = off[i];
  if (t < g.ntries && k.len > 0)
    walk(g.P, g.P.tries[t].root, k, [&](ref_t r) {
      idx[o] = o;
      out[o++] = r;
    });
}

__global__ __launch_bounds__(256) void k_unique_refs(const ref_t *sorted, uint32_t m, uint32_t *flag) {
  uint32_t i = grid_tid();
  if (i >= m) return;
  flag[i] = (i == 0 || sorted[i] != sorted[i - 1]) ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_compact_refs(const ref_t *sorted, const uint32_t *flag,
                                                      const uint32_t *pos, uint32_t m, ref_t *out) {
  uint32_t i = grid_tid();
  if (i >= m) return;
  if (flag[i]) out[pos[i]] = sorted[i];
}

}  // namespace

hipError_t launch_gets(const GetArgs &g, hipStream_t s) {
  if (g.n == 0) return hipSuccess;
  k_gets<<<nblk(g.n), kB, 0, s>>>(g);
  return hipGetLastError();
}

hipError_t launch_gather_values(const uint64_t *addr, const int32_t *size, const uint64_t *off, uint32_t n,
                                uint8_t *out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_gather_values<<<n, 128, 0, s>>>(addr, size, off, n, out);
  return hipGetLastError();
}

#define EXP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e__ = (expr);                                                            \
    if (e__ != hipSuccess) {                                                            \
      set_error(std::string("node export: ") + #expr + ": " + hipGetErrorString(e__));  \
      rc = -1;                                                                          \
      goto done;                                                                        \
    }                                                                                   \
  } while (0)

int export_paths(Forest *f, const BatchDesc &keys, uint8_t **hashes, uint8_t **encs, uint64_t **enc_indexs,
                 uint64_t *num) {
  int rc = 0;
  hipStream_t s = f->stream();
  uint32_t n = keys.n;
  GetArgs g{f->pools(), keys.keys_hexs, keys.keys_indexs, f->ntries() > 1 ? keys.trie_ids : nullptr, n,
            keys.keys_bytes, f->ntries(), nullptr, nullptr, nullptr};
  uint32_t *cnt = nullptr, *off = nullptr, *idx0 = nullptr, *idx1 = nullptr, *flag = nullptr, *pos = nullptr;
  ref_t *r0 = nullptr, *r1 = nullptr, *uniq = nullptr;
  uint64_t *len = nullptr, *eoff = nullptr;
  uint8_t *d_enc = nullptr, *d_hash = nullptr;
  void *temp = nullptr;
  uint32_t total = 0, m = 0;
  uint64_t ebytes = 0;
  int which = 0;
  std::vector<uint64_t> hlen, hoff;
  EXP_TRY(hipMalloc(&cnt, 4ull * n + 4));
  EXP_TRY(hipMalloc(&off, 4ull * n + 4));
  EXP_TRY(hipMalloc(&temp, scan_temp_bytes(n + 1) + 64));
  k_path_count<<<nblk(n), kB, 0, s>>>(g, cnt);
  EXP_TRY(hipGetLastError());
  EXP_TRY(scan_exclusive_u32(cnt, off, n, off + n, temp, s));
  EXP_TRY(hipMemcpyAsync(&total, off + n, 4, hipMemcpyDeviceToHost, s));
  EXP_TRY(hipStreamSynchronize(s));
  if (total == 0) goto done;
  EXP_TRY(hipMalloc(&r0, 4ull * total));
  EXP_TRY(hipMalloc(&r1, 4ull * total));
  EXP_TRY(hipMalloc(&idx0, 4ull * total));
  EXP_TRY(hipMalloc(&idx1, 4ull * total));
  EXP_TRY(hipMalloc(&flag, 4ull * total + 4));
  EXP_TRY(hipMalloc(&pos, 4ull * total + 4));
  EXP_TRY(hipFree(temp));
  temp = nullptr;
  EXP_TRY(hipMalloc(&temp, radix_temp_bytes(total) + scan_temp_bytes(total + 1) + 64));
  k_path_write<<<nblk(n), kB, 0, s>>>(g, off, r0, idx0);
  EXP_TRY(hipGetLastError());
  EXP_TRY(radix_sort_u32(r0, idx0, r1, idx1, total, 0, 32, &which, temp, s));
  {
    ref_t *sorted = which ? r1 : r0;
    k_unique_refs<<<nblk(total), kB, 0, s>>>(sorted, total, flag);
    EXP_TRY(hipGetLastError());
    EXP_TRY(scan_exclusive_u32(flag, pos, total, pos + total, temp, s));
    EXP_TRY(hipMemcpyAsync(&m, pos + total, 4, hipMemcpyDeviceToHost, s));
    EXP_TRY(hipStreamSynchronize(s));
    EXP_TRY(hipMalloc(&uniq, 4ull * m));
    k_compact_refs<<<nblk(total), kB, 0, s>>>(sorted, flag, pos, total, uniq);
    EXP_TRY(hipGetLastError());
  }
  EXP_TRY(hipMalloc(&len, 8ull * m + 8));
  EXP_TRY(hipMalloc(&eoff, 8ull * m + 8));
  EXP_TRY(hc_enc_len(f->pools(), uniq, m, len, s));
  EXP_TRY(scan_exclusive_u64(len, eoff, m, eoff + m, temp, s));
  EXP_TRY(hipMemcpyAsync(&ebytes, eoff + m, 8, hipMemcpyDeviceToHost, s));
  EXP_TRY(hipStreamSynchronize(s));
  EXP_TRY(hipMalloc(&d_enc, ebytes + 8));
  EXP_TRY(hipMalloc(&d_hash, 32ull * m + 8));
  EXP_TRY(hc_enc_write(f->pools(), uniq, m, eoff, d_enc, d_hash, s));
  hlen.resize(m);
  hoff.resize(m);
  *hashes = new uint8_t[32ull * m];
  *encs = new uint8_t[ebytes ? ebytes : 1];
  *enc_indexs = new uint64_t[2ull * m];
  EXP_TRY(hipMemcpyAsync(*hashes, d_hash, 32ull * m, hipMemcpyDeviceToHost, s));
  EXP_TRY(hipMemcpyAsync(*encs, d_enc, ebytes, hipMemcpyDeviceToHost, s));
  EXP_TRY(hipMemcpyAsync(hlen.data(), len, 8ull * m, hipMemcpyDeviceToHost, s));
  EXP_TRY(hipMemcpyAsync(hoff.data(), eoff, 8ull * m, hipMemcpyDeviceToHost, s));
  EXP_TRY(hipStreamSynchronize(s));
  for (uint32_t i = 0; i < m; i++) {
    (*enc_indexs)[2 * i] = hoff[i];
    (*enc_indexs)[2 * i + 1] = hoff[i] + hlen[i] - 1;
  }
  *num = m;
done:
  void *bufs[] = {cnt, off, idx0, idx1, flag, pos, r0, r1, uniq, len, eoff, d_enc, d_hash, temp};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  return rc;
}

namespace {

__global__ __launch_bounds__(256) void k_rw_mark(const uint8_t *flags, uint32_t n, uint32_t *isr, uint32_t *err) {
  uint32_t i = grid_tid();
  if (i >= n) return;
  uint8_t f = flags[i];
  if (f > 1) atomicCAS(err, 0u, (uint32_t)ERR_BAD_KEY);
  isr[i] = f == 0 ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_rw_scatter(const uint8_t *flags, uint32_t n, const int32_t *kidx,
                                                    const int64_t *vidx, const uint64_t *hps, const int32_t *tid,
                                                    const uint32_t *pos, int32_t *rk, int32_t *rtid, int32_t *wk,
                                                    int64_t *wv, uint64_t *whp, int32_t *wtid) {
  uint32_t i = grid_tid();
  if (i >= n) return;
  uint32_t r = pos[i];
  if (flags[i] == 0) {
    rk[2 * r] = kidx[2 * i];
    rk[2 * r + 1] = kidx[2 * i + 1];
    if (tid) rtid[r] = tid[i];
  } else {
    uint32_t w = i - r;
    wk[2 * w] = kidx[2 * i];
    wk[2 * w + 1] = kidx[2 * i + 1];
    wv[2 * w] = vidx[2 * i];
    wv[2 * w + 1] = vidx[2 * i + 1];
    if (hps) whp[w] = hps[i];
    if (tid) wtid[w] = tid[i];
  }
}

}  // namespace

hipError_t rw_split(const uint8_t *flags, uint32_t n, const int32_t *kidx, const int64_t *vidx, const uint64_t *hps,
                    const int32_t *tid, uint32_t *isr, uint32_t *pos, uint32_t *nread, int32_t *rk, int32_t *rtid,
                    int32_t *wk, int64_t *wv, uint64_t *whp, int32_t *wtid, uint32_t *err, void *temp,
                    hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_rw_mark<<<nblk(n), kB, 0, s>>>(flags, n, isr, err);
  GMPT_LAUNCH_CHECK();
  hipError_t e = scan_exclusive_u32(isr, pos, n, nread, temp, s);
  if (e != hipSuccess) return e;
  k_rw_scatter<<<nblk(n), kB, 0, s>>>(flags, n, kidx, vidx, hps, tid, pos, rk, rtid, wk, wv, whp, wtid);
  return hipGetLastError();
}

}  // namespace gmpt
