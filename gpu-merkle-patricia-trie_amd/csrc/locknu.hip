// locknu.hip -- LockNU: latch-coupled concurrent insert (optimistic lock coupling).
//
// Reference: puts_latching_with_valuehp_v2 (mpt-with-compress/include/mpt/
// gpu_mpt.cuh:554-615) launches one warp per key with only lane 0 working
// (puts_latching_v2, gpu_mpt_kernels.cuh:1573-1605) and descends with OLC
// (put_olc_v2, :971-1228; version words node.cuh:359-396, lock.cuh:31-100).
//
// CDNA wavefronts have no independent thread scheduling: a lane that spins on
// a latch held by another lane of its own wavefront can starve it.  Here every
// lane inserts its own key and never waits: a latch is taken with one CAS
// (try-lock); when the CAS fails or a version check fails the lane abandons
// the attempt and restarts from the root in the next trip of a loop that the
// whole wavefront runs together.  A lane that holds a latch always finishes
// its critical section in the same trip, so no lane can block another.
//
// Version word: bit1 = locked, bit0 = obsolete; unlock adds 2 (version + 4
// per write).  Duplicate keys resolve to the highest batch position (last
// writer wins, as the sequential CPU insert does): a value is replaced only by
// a value with a larger order stamp.
//
// Node slots are pre-assigned per key (one branch, two short nodes, one value
// node), so the insert needs no allocator atomics; slots that a key does not
// use stay unreferenced.
#include "locknu.h"
#include "prims.h"

namespace gmpt {

namespace {

constexpr int kB = 256;
__device__ __forceinline__ uint32_t grid_tid() { return blockIdx.x * blockDim.x + threadIdx.x; }
inline unsigned nblk(size_t n) { return (unsigned)((n + kB - 1) / kB); }

// Every word another lane may change during the insert (versions, child
// references, short-node fields, value order stamps) is read with a coherent
// agent-scope load.  Such loads read memory in the order they return, so the
// optimistic read protocol only needs each load of a step to have returned
// before the next step's loads issue: load_order() waits for outstanding
// loads.  An acquire load would also invalidate the CU's caches at every
// version check.
template <typename T>
__device__ __forceinline__ T ld(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void load_order() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
// version read before a node's contents
__device__ __forceinline__ uint64_t ver_read(uint64_t *l) {
  uint64_t v = ld(l);
  load_order();
  return v;
}
// version re-check after the contents were read
__device__ __forceinline__ bool ver_check(uint64_t *l, uint64_t v) {
  load_order();
  return ld(l) == v;
}
__device__ __forceinline__ bool ver_upgrade(uint64_t *l, uint64_t v) {
  uint64_t exp = v;
  return __hip_atomic_compare_exchange_strong(l, &exp, v + 2, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ver_unlock(uint64_t *l) {
  __hip_atomic_fetch_add(l, (uint64_t)2, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ver_unlock_obsolete(uint64_t *l) {
  __hip_atomic_fetch_add(l, (uint64_t)3, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

struct KeyRef {
  const uint8_t *p;
  uint32_t len;
  __device__ __forceinline__ uint32_t nib(uint32_t j) const { return j + 1 == len ? kNibbleT : packed_nibble(p, j); }
};

struct Ctx {
  KeyRef k;
  uint32_t i;       // batch position
  uint32_t trie;
  uint64_t seq;
  uint64_t key_off; // this key's packed bytes in the key arena
  uint64_t key_cap; // key arena bytes in use (bounds for optimistic reads)
};

__device__ __forceinline__ void init_value(const NuArgs &a, const Ctx &c, uint32_t vid) {
  Value &V = a.P.value[vid];
  int64_t vs = a.vidx[2 * c.i], ve = a.vidx[2 * c.i + 1];
  V.addr = a.vbase + (uint64_t)vs;
  V.hp = a.hps ? a.hps[c.i] : 0ull;
  V.size = (uint32_t)(ve - vs + 1);
  V.pad = 0;
  V.seq = c.seq;
}

__device__ __forceinline__ void init_short(Short &S, uint64_t key_off, uint32_t key_start, uint32_t key_len,
                                           uint32_t depth, uint8_t term, ref_t child) {
  uint64_t *w = (uint64_t *)&S;
#pragma unroll
  for (int q = 0; q < 8; q++) w[q] = 0;
  S.key_off = key_off;
  S.child = child;
  S.key_start = (uint16_t)key_start;
  S.key_len = (uint16_t)key_len;
  S.depth = (uint16_t)depth;
  S.term = term;
  S.dirty = 1;
}

// New leaf for the key's nibbles [start, len): a short node over a value node,
// or the value node alone when nothing is left (key ends at a branch slot 16).
__device__ __forceinline__ ref_t make_leaf(const NuArgs &a, const Ctx &c, uint32_t start, bool *made_short) {
  uint32_t vid = a.base_value + c.i;
  init_value(a, c, vid);
  ref_t v = make_ref(T_VALUE, vid);
  *made_short = start < c.k.len;
  if (!*made_short) return v;
  uint32_t sid = a.base_short + 2 * c.i;
  init_short(a.P.shrt[sid], c.key_off, start, c.k.len - start, start, 1, v);
  return make_ref(T_SHORT, sid);
}

enum { ATT_DONE = 0, ATT_RESTART = 1, ATT_FAIL = 2 };

// One optimistic attempt.  Returns ATT_DONE, or ATT_RESTART when a version
// check or latch upgrade failed (nothing was modified).  Path nodes are marked
// dirty on the way down; the nodes a finished insert modified or created are
// reported through mk for the caller to mark.
struct Marks {
  ref_t existing[2];   // existing nodes whose encoding changes
  ref_t fresh[3];      // new short/branch nodes
  int ne, nf;
};

__device__ int attempt(const NuArgs &a, const Ctx &c, Marks &mk, DirtyStage &ds) {
  const Pools &P = a.P;
  mk.ne = mk.nf = 0;
  TrieHdr *H = &P.tries[c.trie];
  uint64_t *plock = &H->lock;
  uint64_t pver = ver_read(plock);
  if (pver & 3) return ATT_RESTART;
  ref_t *pslot = &H->root;
  ref_t parent = REF_NULL;
  ref_t node = ld(pslot);
  if (!ver_check(plock, pver)) return ATT_RESTART;
  uint32_t pos = 0;
  for (int guard = 0; guard < 2 * kMaxKeyNibbles + 4; guard++) {
    uint32_t ty = ref_type(node), idx = ref_idx(node);
    if (node == REF_NULL) {
      if (!ver_upgrade(plock, pver)) return ATT_RESTART;
      bool ms;
      ref_t leaf = make_leaf(a, c, pos, &ms);
      __hip_atomic_store(pslot, leaf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      ver_unlock(plock);
      if (ms) mk.fresh[mk.nf++] = leaf;
      if (parent != REF_NULL) mk.existing[mk.ne++] = parent;
      return ATT_DONE;
    }
    if (ty == T_BRANCH) {
      if (idx >= P.cap_branch) return ATT_RESTART;
      Branch *B = &P.branch[idx];
      uint64_t v = ver_read(&B->lock);
      if (v & 3) return ATT_RESTART;
      if (!ver_check(plock, pver)) return ATT_RESTART;
      if (pos >= c.k.len) return ATT_RESTART;
      uint32_t nib = c.k.nib(pos);
      ref_t child = ld(&B->child[nib]);
      if (!ver_check(&B->lock, v)) return ATT_RESTART;
      // on the key's path: re-hash (a mark left by an abandoned attempt only
      // costs a redundant re-hash)
      ds.mark(P, node);
      parent = node;
      plock = &B->lock;
      pver = v;
      pslot = &B->child[nib];
      node = child;
      pos += 1;
      continue;
    }
    if (ty == T_SHORT) {
      if (idx >= P.cap_short) return ATT_RESTART;
      Short *S = &P.shrt[idx];
      uint64_t v = ver_read(&S->lock);
      if (v & 3) return ATT_RESTART;
      if (!ver_check(plock, pver)) return ATT_RESTART;
      // fields read as the two 64-bit words that hold them (offsets 40, 48)
      uint64_t koff = ld(&S->key_off);
      uint64_t w40 = ld((const uint64_t *)((const uint8_t *)S + 40));
      uint64_t w48 = ld((const uint64_t *)((const uint8_t *)S + 48));
      ref_t child = (ref_t)(w40 & 0xffffffffu);
      uint32_t kstart = (uint32_t)((w40 >> 32) & 0xffffu), klen = (uint32_t)(w40 >> 48);
      uint32_t depth = (uint32_t)(w48 & 0xffffu), term = (uint32_t)((w48 >> 24) & 0xffu);
      if (klen == 0 || klen > kMaxKeyNibbles || depth != pos || koff + (kstart + klen) / 2 + 1 > c.key_cap)
        return ATT_RESTART;
      const uint8_t *sk = P.keys + koff;
      uint32_t p = 0;
      while (p < klen && pos + p < c.k.len) {
        uint32_t sn = (term && p + 1 == klen) ? kNibbleT : packed_nibble(sk, kstart + p);
        if (sn != c.k.nib(pos + p)) break;
        p++;
      }
      if (!ver_check(&S->lock, v)) return ATT_RESTART;
      if (p == klen) {
        if (ref_type(child) == T_VALUE) {
          // the key exists: swap the value if this write is later.  A lane
          // whose write is already superseded finishes without the latch.
          uint32_t cv = ref_idx(child);
          if (cv < P.cap_value) {
            uint64_t cur = ld(&P.value[cv].seq);
            if (!ver_check(&S->lock, v)) return ATT_RESTART;
            if (cur >= c.seq) return ATT_DONE;
          }
          if (!ver_upgrade(&S->lock, v)) return ATT_RESTART;
          uint32_t vi = ref_idx(child);
          bool newer = vi >= P.cap_value || P.value[vi].seq < c.seq;
          if (newer) {
            uint32_t vid = a.base_value + c.i;
            init_value(a, c, vid);
            __hip_atomic_store(&S->child, make_ref(T_VALUE, vid), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
          ver_unlock(&S->lock);
          mk.existing[mk.ne++] = node;
          return ATT_DONE;
        }
        ds.mark(P, node);
        parent = node;
        plock = &S->lock;
        pver = v;
        pslot = &S->child;
        node = child;
        pos += klen;
        continue;
      }
      if (pos + p >= c.k.len) return ATT_RESTART;
      // split S at p: a new branch at depth pos + p
      uint32_t d = pos + p;
      uint32_t snib = (term && p + 1 == klen) ? kNibbleT : packed_nibble(sk, kstart + p);
      uint32_t knib = c.k.nib(d);
      if (p == 0 && !ver_upgrade(plock, pver)) return ATT_RESTART;
      if (!ver_upgrade(&S->lock, v)) {
        if (p == 0) ver_unlock(plock);
        return ATT_RESTART;
      }
      uint32_t bid = a.base_branch + c.i;
      Branch &Bn = P.branch[bid];
      {
        uint64_t *w = (uint64_t *)&Bn;
#pragma unroll
        for (int q = 0; q < 16; q++) w[q] = 0;
      }
      Bn.depth = (uint16_t)d;
      Bn.dirty = 1;
      ref_t rest = child;
      if (p + 1 < klen) {
        uint32_t rid = a.base_short + 2 * c.i + 1;
        init_short(P.shrt[rid], koff, kstart + p + 1, klen - p - 1, d + 1, (uint8_t)term, child);
        rest = make_ref(T_SHORT, rid);
        mk.fresh[mk.nf++] = rest;
      }
      Bn.child[snib] = rest;
      bool ms;
      ref_t leaf = make_leaf(a, c, d + 1, &ms);
      Bn.child[knib] = leaf;
      if (ms) mk.fresh[mk.nf++] = leaf;
      ref_t bref = make_ref(T_BRANCH, bid);
      mk.fresh[mk.nf++] = bref;
      if (p == 0) {
        __hip_atomic_store(pslot, bref, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ver_unlock_obsolete(&S->lock);
        ver_unlock(plock);
        if (parent != REF_NULL) mk.existing[mk.ne++] = parent;
      } else {
        S->key_len = (uint16_t)p;
        S->term = 0;
        __hip_atomic_store(&S->child, bref, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ver_unlock(&S->lock);
        mk.existing[mk.ne++] = node;
      }
      return ATT_DONE;
    }
    // a value node in a branch's slot 16: the key exists
    if (parent == REF_NULL || ref_type(parent) != T_BRANCH) return ATT_RESTART;
    if (!ver_upgrade(plock, pver)) return ATT_RESTART;
    uint32_t vi = ref_idx(node);
    if (vi >= P.cap_value || P.value[vi].seq < c.seq) {
      uint32_t vid = a.base_value + c.i;
      init_value(a, c, vid);
      __hip_atomic_store(pslot, make_ref(T_VALUE, vid), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    ver_unlock(plock);
    mk.existing[mk.ne++] = parent;
    return ATT_DONE;
  }
  return ATT_RESTART;
}

// Lane j inserts the j-th surviving write: batch position u_idx[j] (the last
// write of each distinct key, scal[SC_M_UNIQ] of them).
constexpr uint32_t kDirtyStage = 4096;  // staged dirty-list entries per block

__global__ __launch_bounds__(256) void k_lock_insert(NuArgs a, uint64_t key_cap, uint32_t max_trips) {
  __shared__ ref_t dbuf[kDirtyStage];
  __shared__ uint32_t dcnt;
  if (threadIdx.x == 0) dcnt = 0;
  __syncthreads();
  DirtyStage ds{dbuf, &dcnt, kDirtyStage};
  const uint32_t j = grid_tid();
  const bool active = j < a.scal[SC_M_UNIQ];
  const uint32_t i = active ? a.u_idx[j] : 0u;
  Ctx c;
  c.i = i;
  c.key_cap = key_cap;
  if (active) {
    const uint64_t kr = a.kref[i];
    c.key_off = a.kbase + (kr & ((1ull << 48) - 1));
    c.k.p = a.P.keys + c.key_off;
    c.k.len = (uint32_t)(kr >> 48);
    c.trie = a.tid ? (uint32_t)a.tid[i] : 0u;
    c.seq = a.seq_base + i;
  }
  bool done = !active;
  uint32_t restarts = 0;
  Marks mk;
  mk.ne = mk.nf = 0;
  for (uint32_t trip = 0; trip < max_trips; trip++) {
    int res = ATT_RESTART;
    if (!done) {
      res = attempt(a, c, mk, ds);
      if (res == ATT_DONE)
        done = true;
      else
        restarts++;
    }
    if (res == ATT_DONE) {
      // the modified existing nodes, then the new nodes
      for (int q = 0; q < mk.ne; q++) ds.mark(a.P, mk.existing[q]);
      for (int q = 0; q < mk.nf; q++) ds.push(a.P, mk.fresh[q]);
    }
    if (__ballot(!done) == 0) break;
  }
  dirty_flush(a.P, ds);
  if (active && !done) raise_error(a.P.cnt, ERR_INTERNAL, 20);
  // one atomic pair per wave (per-lane atomics on one counter serialise)
  uint32_t fails = restarts ? 1u : 0u;
  for (int off = 32; off > 0; off >>= 1) {
    restarts += (uint32_t)__shfl_xor((int)restarts, off);
    fails += (uint32_t)__shfl_xor((int)fails, off);
  }
  if (__lane_id() == 0 && restarts) {
    atomicAdd(&a.P.cnt->n_restarts, restarts);
    atomicAdd(&a.P.cnt->n_lock_fail, fails);
  }
}

__global__ void k_bump(Counters *cnt, uint32_t nb, uint32_t ns, uint32_t nv) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    cnt->n_branch += nb;
    cnt->n_short += ns;
    cnt->n_value += nv;
  }
}

}  // namespace

hipError_t lock_insert(const NuArgs &a, uint64_t key_cap, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  k_lock_insert<<<nblk(a.n), kB, 0, s>>>(a, key_cap, 1u << 20);
  GMPT_LAUNCH_CHECK();
  k_bump<<<1, 64, 0, s>>>(a.P.cnt, a.n, 2 * a.n, a.n);
  return hipGetLastError();
}

}  // namespace gmpt
