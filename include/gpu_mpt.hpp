// gpu_mpt.hpp -- GpuMPT::Compress::MPT, the reference's host class
// (mpt-with-compress/include/mpt/gpu_mpt.cuh:18-192), as a header-only C++
// wrapper over the C-ABI in gmpt_amd.h.  Same method names, argument meaning
// and return shapes; link with libgmpt_amd.so.
//
// Differences, each deliberate:
//  - Errors throw gmpt_error (the reference calls exit(EXIT_FAILURE)).
//  - Inputs are copied into the trie's own HBM pools; the caller may free its
//    buffers after a call (the reference keeps pointers into per-call device
//    copies that it never frees).
//  - Duplicate keys in a batch resolve to the last occurrence in both
//    PhaseNU and LockNU, and PhaseNU overwrites existing keys (the CPU MPT's
//    semantics; the reference GPU paths keep the first or an arbitrary one).
//  - The returned hash-target handle (Node**) is the trie's device-side dirty
//    list; it is only meaningful to hash_onepass_v2 of the same MPT.
#pragma once

#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "gmpt_amd.h"

namespace GpuMPT {
namespace Compress {

struct Node;  // opaque: node handles live in device memory

class gmpt_error : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

class MPT {
 public:
  explicit MPT(int device = 0) {
    check(gmpt_forest_create(device, 1, &f_));
  }
  ~MPT() { gmpt_forest_destroy(f_); }
  MPT(const MPT &) = delete;
  MPT &operator=(const MPT &) = delete;

  // gpu_mpt.cuh:95-97, 1262-1399 (PhaseNU)
  std::tuple<Node **, int> puts_2phase_with_valuehp(const uint8_t *keys_hexs, int *keys_indexs,
                                                    const uint8_t *values_bytes, int64_t *values_indexs,
                                                    const uint8_t **values_hps, int n) {
    gmpt_batch b = batch(keys_hexs, keys_indexs, values_bytes, values_indexs, values_hps, n);
    void *t = nullptr;
    int nt = 0;
    check(gmpt_puts_2phase(f_, &b, &t, &nt));
    return {static_cast<Node **>(t), nt};
  }

  // gpu_mpt.cuh:1249-1260: host pointers into values_bytes are recorded
  std::tuple<Node **, int> puts_2phase(const uint8_t *keys_hexs, int *keys_indexs, const uint8_t *values_bytes,
                                       int64_t *values_indexs, int n) {
    std::vector<const uint8_t *> hps(n > 0 ? n : 1);
    for (int i = 0; i < n; i++) hps[i] = values_bytes + values_indexs[2 * i];
    return puts_2phase_with_valuehp(keys_hexs, keys_indexs, values_bytes, values_indexs, hps.data(), n);
  }

  // gpu_mpt.cuh:69-71, 554-615 (LockNU)
  std::tuple<Node **, int> puts_latching_with_valuehp_v2(const uint8_t *keys_hexs, int *keys_indexs,
                                                         const uint8_t *values_bytes, int64_t *values_indexs,
                                                         const uint8_t **values_hps, int n) {
    gmpt_batch b = batch(keys_hexs, keys_indexs, values_bytes, values_indexs, values_hps, n);
    void *t = nullptr;
    int nt = 0;
    check(gmpt_puts_latching(f_, &b, &t, &nt));
    return {static_cast<Node **>(t), nt};
  }

  // gpu_mpt.cuh:105-110, 1609-1713 (mixed reads and writes, PhaseNU writes);
  // reads see the trie before the batch's writes and come back in batch order
  std::tuple<Node **, int> puts_2phase_with_valuehp_with_read(const uint8_t *keys_hexs, int *keys_indexs,
                                                              int &read_num, const uint8_t *rw_flags,
                                                              const uint8_t *values_bytes, int64_t *values_indexs,
                                                              const uint8_t **values_hps, int n,
                                                              const uint8_t **read_values_hps,
                                                              int *read_values_sizes) {
    return with_read(0, keys_hexs, keys_indexs, read_num, rw_flags, values_bytes, values_indexs, values_hps, n,
                     read_values_hps, read_values_sizes);
  }

  // gpu_mpt.cuh:74-78, 684-765 (mixed reads and writes, LockNU writes)
  std::tuple<Node **, int> puts_latching_with_valuehp_v2_with_read(const uint8_t *keys_hexs, int *keys_indexs,
                                                                   int &read_num, const uint8_t *rw_flags,
                                                                   const uint8_t *values_bytes,
                                                                   int64_t *values_indexs, const uint8_t **values_hps,
                                                                   int n, const uint8_t **read_values_hps,
                                                                   int *read_values_sizes) {
    return with_read(1, keys_hexs, keys_indexs, read_num, rw_flags, values_bytes, values_indexs, values_hps, n,
                     read_values_hps, read_values_sizes);
  }

  // gpu_mpt.cuh:131, 1162-1191 (PhaseHC)
  void hash_onepass_v2(Node **d_hash_nodes, int n) { check(gmpt_hash_onepass(f_, d_hash_nodes, n)); }

  // gpu_mpt.cuh:139-140, 984-1036
  void gets_parallel(const uint8_t *keys_hexs, int *keys_indexs, int n, const uint8_t **values_hps,
                     int *values_sizes) const {
    gmpt_batch b = batch(keys_hexs, keys_indexs, nullptr, nullptr, nullptr, n);
    check(gmpt_gets(f_, &b, values_hps, values_sizes));
  }

  // gpu_mpt.cuh:146-149, 1039-1122
  void flush_dirty_nodes(const uint8_t *keys_hexs, int *keys_indexs, int n_keys, const uint8_t *&hashs,
                         const uint8_t *&encs, const unsigned long long *&encs_indexs, unsigned long long &n_nodes) {
    gmpt_batch b = batch(keys_hexs, keys_indexs, nullptr, nullptr, nullptr, n_keys);
    const uint64_t *ix = nullptr;
    uint64_t num = 0;
    check(gmpt_export_nodes(f_, &b, &hashs, &encs, &ix, &num));
    encs_indexs = reinterpret_cast<const unsigned long long *>(ix);
    n_nodes = num;
  }

  // gpu_mpt.cuh:151, 1217-1240: a new[] 32-byte buffer owned by the caller,
  // nullptr / 0 for an empty trie
  void get_root_hash(const uint8_t *&hash, int &hash_size) const {
    uint8_t *h = new uint8_t[32]();
    int hs = 0;
    int r = gmpt_get_root_hash(f_, 0, h, &hs);
    if (r != 0) {
      delete[] h;
      check(r);
    }
    if (hs == 0) {
      delete[] h;
      h = nullptr;
    }
    hash = h;
    hash_size = hs;
  }

  // gpu_mpt.cuh:152, 1242-1247
  std::tuple<const uint8_t *, int> get_root_hash() const {
    const uint8_t *h;
    int hs;
    get_root_hash(h, hs);
    return {h, hs};
  }

  gmpt_forest *handle() const { return f_; }

 private:
  std::tuple<Node **, int> with_read(int mode, const uint8_t *kh, int *ki, int &read_num, const uint8_t *flags,
                                     const uint8_t *vb, int64_t *vi, const uint8_t **hps, int n,
                                     const uint8_t **rhps, int *rsizes) {
    gmpt_batch b = batch(kh, ki, vb, vi, hps, n);
    void *t = nullptr;
    int nt = 0;
    check(gmpt_puts_with_read(f_, &b, flags, mode, &read_num, rhps, rsizes, &t, &nt));
    return {static_cast<Node **>(t), nt};
  }
  static void check(int rc) {
    if (rc != 0) throw gmpt_error(gmpt_last_error());
  }
  static gmpt_batch batch(const uint8_t *kh, const int *ki, const uint8_t *vb, const int64_t *vi,
                          const uint8_t **hps, int n) {
    gmpt_batch b{};
    b.keys_hexs = kh;
    b.keys_indexs = ki;
    b.values_bytes = vb;
    b.values_indexs = vi;
    b.values_hps = hps;
    b.n = n;
    return b;
  }
  gmpt_forest *f_ = nullptr;
};

}  // namespace Compress
}  // namespace GpuMPT
