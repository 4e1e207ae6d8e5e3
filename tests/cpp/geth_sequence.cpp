// geth_sequence.cpp -- replays go-ethereum's libgmpt call sequence
// (mpt-with-compress/lib/libgmpt.cu:25-331; case_study.sh:21-29) and the
// GpuMPT::Compress::MPT class calls (gpu_mpt.cuh:69-152) through the C-ABI
// headers only, the way a cgo or C++ caller links libgmpt_amd.so.
//
// Input file (little-endian), written by tests/test_cabi_harness.py:
//   u32 nblocks; per block, per trie 0..2:
//     u32 n; u64 key_bytes; u64 value_bytes;
//     key_bytes of hex keys; n*2 int32 key index pairs;
//     value_bytes of values; n*2 int64 value index pairs
// Output lines (the test checks them against the oracle):
//   build <block> <trie> <root hex>       build_mpt_2phase (even blocks) / build_mpt_olc (odd)
//   nodes <num> <hash hex> <enc hex>...   get_all_nodes(STATE_TRIE, last state batch), one line per node
//   insert <block> <root hex>             init_mpt + insert_mpt_2phase / insert_mpt_olc of the state batches
//   class <root hex>                      MPT class: PhaseNU then LockNU over the state batches
//   gets <ok>                             gets_parallel returns the caller's host pointers and sizes
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "gmpt_amd.h"
#include "gpu_mpt.hpp"
#include "libgmpt.h"

namespace {

struct Batch {
  std::vector<uint8_t> keys;
  std::vector<int> kidx;
  std::vector<uint8_t> vals;
  std::vector<int64_t> vidx;
  int n = 0;
};

bool rd(FILE *f, void *p, size_t bytes) { return bytes == 0 || fread(p, 1, bytes, f) == bytes; }

std::string hex(const uint8_t *p, size_t n) {
  static const char *d = "0123456789abcdef";
  std::string s;
  for (size_t i = 0; i < n; i++) {
    s += d[p[i] >> 4];
    s += d[p[i] & 15];
  }
  return s;
}

int die(const char *what) {
  fprintf(stderr, "geth_sequence: %s: %s\n", what, gmpt_last_error());
  return 1;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: geth_sequence <batches.bin>\n");
    return 2;
  }
  FILE *f = fopen(argv[1], "rb");
  if (!f) return die("open input");
  uint32_t nblocks = 0;
  if (!rd(f, &nblocks, 4)) return die("read header");
  std::vector<Batch> b(3ull * nblocks);
  for (auto &x : b) {
    uint32_t n;
    uint64_t kb, vb;
    if (!rd(f, &n, 4) || !rd(f, &kb, 8) || !rd(f, &vb, 8)) return die("read batch header");
    x.n = (int)n;
    x.keys.resize(kb + 8);
    x.kidx.resize(2ull * n + 2);
    x.vals.resize(vb + 8);
    x.vidx.resize(2ull * n + 2);
    if (!rd(f, x.keys.data(), kb) || !rd(f, x.kidx.data(), 8ull * n) || !rd(f, x.vals.data(), vb) ||
        !rd(f, x.vidx.data(), 16ull * n))
      return die("read batch");
  }
  fclose(f);

  // 1. preprocess + per-block builds of the three tries
  if (preprocess() != 1) return die("preprocess");
  for (uint32_t blk = 0; blk < nblocks; blk++)
    for (int t = 0; t < 3; t++) {
      Batch &x = b[3 * blk + t];
      auto fn = blk % 2 == 0 ? build_mpt_2phase : build_mpt_olc;
      const uint8_t *root = fn((TrieType)t, x.keys.data(), x.kidx.data(), x.vals.data(), x.vidx.data(), nullptr, x.n);
      if (!root) return die("build_mpt");
      printf("build %u %d %s\n", blk, t, hex(root, 32).c_str());
      delete[] root;
    }

  // 2. node export of the state trie along the last block's keys
  {
    Batch &x = b[3 * (nblocks - 1)];
    nodeset ns = get_all_nodes(STATE_TRIE, x.keys.data(), x.kidx.data(), x.n);
    if (ns.num == 0) return die("get_all_nodes");
    for (unsigned long long i = 0; i < ns.num; i++) {
      unsigned long long s = ns.encs_indexs[2 * i], e = ns.encs_indexs[2 * i + 1];
      printf("nodes %llu %s %s\n", ns.num, hex(ns.hashs + 32 * i, 32).c_str(), hex(ns.encs + s, e - s + 1).c_str());
    }
    gmpt_free_host(ns.hashs);
    gmpt_free_host(ns.encs);
    gmpt_free_host(ns.encs_indexs);
  }

  // 3. incremental handle API
  void *m = init_mpt();
  if (!m) return die("init_mpt");
  for (uint32_t blk = 0; blk < nblocks; blk++) {
    Batch &x = b[3 * blk];
    auto fn = blk % 2 == 0 ? insert_mpt_2phase : insert_mpt_olc;
    const uint8_t *root = fn(m, x.keys.data(), x.kidx.data(), x.vals.data(), x.vidx.data(), nullptr, x.n);
    if (!root) return die("insert_mpt");
    printf("insert %u %s\n", blk, hex(root, 32).c_str());
    delete[] root;
  }
  free_mpt(m);

  // 4. the reference class, with caller host pointers for gets
  try {
    GpuMPT::Compress::MPT mpt;
    std::vector<std::vector<const uint8_t *>> hps(nblocks);
    for (uint32_t blk = 0; blk < nblocks; blk++) {
      Batch &x = b[3 * blk];
      hps[blk].resize(x.n + 1);
      for (int i = 0; i < x.n; i++) hps[blk][i] = x.vals.data() + x.vidx[2 * i];
      auto r = blk % 2 == 0 ? mpt.puts_2phase_with_valuehp(x.keys.data(), x.kidx.data(), x.vals.data(),
                                                           x.vidx.data(), hps[blk].data(), x.n)
                            : mpt.puts_latching_with_valuehp_v2(x.keys.data(), x.kidx.data(), x.vals.data(),
                                                                x.vidx.data(), hps[blk].data(), x.n);
      mpt.hash_onepass_v2(std::get<0>(r), std::get<1>(r));
    }
    auto [root, size] = mpt.get_root_hash();
    printf("class %s\n", hex(root, size).c_str());
    delete[] root;
    Batch &x = b[3 * (nblocks - 1)];
    std::vector<const uint8_t *> got(x.n + 1);
    std::vector<int> sizes(x.n + 1);
    mpt.gets_parallel(x.keys.data(), x.kidx.data(), x.n, got.data(), sizes.data());
    int ok = 1;
    for (int i = 0; i < x.n; i++)
      if (got[i] != hps[nblocks - 1][i] || sizes[i] != (int)(x.vidx[2 * i + 1] - x.vidx[2 * i] + 1)) ok = 0;
    printf("gets %d\n", ok);
  } catch (const std::exception &e) {
    fprintf(stderr, "geth_sequence: %s\n", e.what());
    return 1;
  }
  return 0;
}
