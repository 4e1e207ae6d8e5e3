// common.h -- node layout in HBM, node references and shared device helpers.
//
// The trie lives in typed pools in device memory.  A node is named by a 32-bit
// tagged reference: bits 31..30 hold the pool (branch / short / value), bits
// 29..0 the slot in that pool.  Pools are plain arrays, so a pool can be
// regrown by copy without rewriting any reference.
//
// Node kinds follow the reference's GPU node model
// (mpt-with-compress/include/mpt/node.cuh:336-589):
//   Branch  <- FullNode  (17 children, 32-byte hash/encoding buffer)
//   Short   <- ShortNode (key + one child; a leaf when the child is a value)
//   Value   <- ValueNode (device value bytes + caller's host pointer)
// Keys are not copied per node: every inserted key is stored once, packed two
// nibbles per byte, in a key arena, and a Short node names a nibble range of
// one such key (the reference points ShortNode::key into the batch's hex keys).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace gmpt {

typedef uint32_t ref_t;

enum : uint32_t { REF_NULL = 0u, T_BRANCH = 1u, T_SHORT = 2u, T_VALUE = 3u };

__host__ __device__ __forceinline__ ref_t make_ref(uint32_t type, uint32_t idx) { return (type << 30) | idx; }
__host__ __device__ __forceinline__ uint32_t ref_type(ref_t r) { return r >> 30; }
__host__ __device__ __forceinline__ uint32_t ref_idx(ref_t r) { return r & 0x3fffffffu; }

constexpr uint32_t kNibbleT = 16;  // hex terminator
constexpr int kMaxKeyNibbles = 254;  // hex key length limit (incl. terminator)

// 128 bytes: one L2 line.
struct Branch {
  uint8_t hash[32];      // Keccak-256 digest, or the raw encoding when it is < 32 B
  uint8_t hash_size;     // 32, or the raw encoding length (< 32)
  uint8_t pad0;
  uint16_t depth;        // nibble position this branch dispatches on
  uint32_t dirty;        // 1 while the node waits for re-hashing
  uint64_t lock;         // OLC version word: bit1 locked, bit0 obsolete, +4 per write
  ref_t child[17];       // slots 0..15 by nibble, 16 for a key ending here
  uint32_t pad1;
  uint8_t pad2[8];
};
static_assert(sizeof(Branch) == 128, "Branch layout");
// A parent reading a child's digest and size touches the child's first 64
// bytes only; the 16 nibble slots load as four 16-byte words.
static_assert(offsetof(Branch, hash_size) < 64 && offsetof(Branch, child) % 16 == 0, "Branch layout");

// 64 bytes.
struct Short {
  uint8_t hash[32];
  uint64_t key_off;      // byte offset of the source key's packed nibbles in the key arena
  ref_t child;           // BRANCH (extension) or VALUE (leaf) reference
  uint16_t key_start;    // first nibble of this node's key inside the source key
  uint16_t key_len;      // nibbles in this node's key, terminator included when term = 1
  uint16_t depth;        // absolute nibble position where this node's key starts
  uint8_t hash_size;
  uint8_t term;          // 1 when the key ends with the terminator (a leaf key)
  uint32_t dirty;
  uint64_t lock;
};
static_assert(sizeof(Short) == 64, "Short layout");

// 32 bytes.
struct Value {
  uint64_t addr;         // device address of the value bytes
  uint64_t hp;           // caller-supplied host pointer returned by gets (0 if none)
  uint32_t size;
  uint32_t pad;
  uint64_t seq;          // batch order stamp (LockNU keeps the highest)
};
static_assert(sizeof(Value) == 32, "Value layout");

// One trie of a forest.
struct TrieHdr {
  ref_t root;
  uint32_t root_hash_size;
  uint64_t lock;         // OLC version word guarding `root`
  uint8_t root_hash[32];
};
// k_hc_roots writes root_hash as 8-byte words and Counters::root0 as 4-byte words
static_assert(offsetof(TrieHdr, root_hash) == 16 && sizeof(TrieHdr) % 8 == 0, "TrieHdr layout");

// Device-side counters, read back by the host after each operation.
struct Counters {
  uint32_t n_branch, n_short, n_value, n_dirty;
  uint64_t key_used, value_used;
  uint32_t err;          // first error code raised by a kernel (0 = none)
  uint32_t err_arg;
  uint32_t flag_fallback;  // PhaseNU: a sort window tie between distinct keys
  uint32_t n_restarts;   // LockNU: optimistic restarts
  uint32_t n_lock_fail;  // LockNU: keys that restarted at least once
  uint32_t graph_bad;    // a replayed pass found its batch unlike the captured one: hash kernels skip
  unsigned long long n_perms;    // PhaseHC: Keccak-f permutations run
  unsigned long long enc_bytes;  // PhaseHC: bytes of node encodings produced
  unsigned long long leaf_perms; // PhaseHC: permutations of the leaf level
  unsigned long long leaf_bytes; // PhaseHC: encoding bytes of the leaf level
  uint32_t scal[20];             // the current insert's device scalars (PhaseNU / LockNU)
  uint32_t root0_size;           // trie 0's root as the last hash pass left it (k_hc_roots), so that
  uint8_t root0[32];             // the pass's closing counter read also brings the root to the host
};
static_assert(offsetof(Counters, root0) % 4 == 0, "Counters layout");

enum : uint32_t {
  ERR_NONE = 0,
  ERR_BAD_KEY = 1,        // key without terminator, interior terminator, nibble > 15, or too long
  ERR_POOL = 2,           // a pool ran out (the host sized it too small)
  ERR_INTERNAL = 3,
};
// err_arg of the ERR_POOL raised when a batch's packed keys exceed the key
// arena (the host then grows the arena to the batch's packed total)
constexpr uint32_t kErrArgKeyArena = 6;

// Device pool views passed to kernels by value.
struct Pools {
  Branch *branch;
  Short *shrt;
  Value *value;
  uint8_t *keys;         // key arena
  TrieHdr *tries;
  ref_t *dirty;          // dirty list (hash targets)
  Counters *cnt;
  uint32_t cap_branch, cap_short, cap_value, cap_dirty;
};

__device__ __forceinline__ void raise_error(Counters *c, uint32_t code, uint32_t arg) {
  if (atomicCAS(&c->err, 0u, code) == 0u) c->err_arg = arg;
}

// Nibble j of a packed key (high nibble first, two per byte).
__device__ __forceinline__ uint32_t packed_nibble(const uint8_t *k, uint32_t j) {
  uint32_t b = k[j >> 1];
  return (j & 1) ? (b & 15u) : (b >> 4);
}

__device__ __forceinline__ uint32_t *node_dirty_flag(const Pools &P, ref_t r) {
  uint32_t t = ref_type(r), i = ref_idx(r);
  return t == T_BRANCH ? &P.branch[i].dirty : &P.shrt[i].dirty;
}

// Block-staged dirty-list appends: lanes stage the nodes they flip in LDS
// (an LDS atomic each) and the block appends them to the global list with one
// global atomic in dirty_flush().  Appending straight to the list costs a
// global atomic on one counter per wave and mark, which serialises when many
// waves mark paths.  A full stage falls back to direct appends.
struct DirtyStage {
  ref_t *buf;     // LDS
  uint32_t *cnt;  // LDS
  uint32_t cap;
  __device__ __forceinline__ void push(const Pools &P, ref_t r) {
    uint32_t pos = atomicAdd(cnt, 1u);
    if (pos < cap) {
      buf[pos] = r;
      return;
    }
    uint32_t slot = atomicAdd(&P.cnt->n_dirty, 1u);
    if (slot < P.cap_dirty)
      P.dirty[slot] = r;
    else
      raise_error(P.cnt, ERR_POOL, 4);
  }
  // marks an existing node: the lane that flips its flag stages it
  __device__ __forceinline__ void mark(const Pools &P, ref_t r) {
    uint32_t *f = node_dirty_flag(P, r);
    if (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0 && atomicExch(f, 1u) == 0u)
      push(P, r);
  }
};

// Every thread of the block calls this once, after its last push.
__device__ __forceinline__ void dirty_flush(const Pools &P, const DirtyStage &ds) {
  __shared__ uint32_t base;
  __syncthreads();
  const uint32_t n = min(*ds.cnt, ds.cap);
  if (threadIdx.x == 0) base = n ? atomicAdd(&P.cnt->n_dirty, n) : 0u;
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
    uint32_t slot = base + k;
    if (slot < P.cap_dirty)
      P.dirty[slot] = ds.buf[k];
    else
      raise_error(P.cnt, ERR_POOL, 4);
  }
}

}  // namespace gmpt

#define GMPT_LAUNCH_CHECK()                                  \
  do {                                                       \
    hipError_t e_ = hipGetLastError();                       \
    if (e_ != hipSuccess) return e_;                         \
  } while (0)
