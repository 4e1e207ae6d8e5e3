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
// Writes are combined first (one lane per distinct key), so whether a key
// exists is fixed before the insert starts: a read-only descent classifies
// every key.  An existing key's value node is rewritten in place (only that
// key refers to it); a new key gets node slots by its rank among the new keys
// (one branch, two short nodes, one value node), so the insert needs no
// allocator atomics and a batch of overwrites allocates nothing.
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
// loads (a version read, then the contents it guards, then the re-check).
// An acquire load would also invalidate the L2 of the CU's die at every
// version check.  Key bytes in the arena never change while the insert runs
// and are read with plain loads.
template <typename T>
__device__ __forceinline__ T ld(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void load_order() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
// Writes that other lanes may read while the insert runs are agent-scope
// stores too: they are written through to the point of coherence, so once a
// lane's stores have completed (store_order() waits for them) every other
// lane's coherent loads see them.  Publishing a node therefore takes: the
// node's stores, store_order(), the reference store, store_order(), the
// unlock.  Release and acquire orderings would instead write back and
// invalidate the whole L2 of the CU's die at every latch.
template <typename T, typename U>
__device__ __forceinline__ void st(T *p, U v) {
  __hip_atomic_store(p, (T)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_order() { load_order(); }
// The latch holder reads nothing it did not validate before the upgrade, and
// stores only after the upgrade returned, so the upgrade needs no ordering.
__device__ __forceinline__ bool ver_upgrade(uint64_t *l, uint64_t v) {
  uint64_t exp = v;
  return __hip_atomic_compare_exchange_strong(l, &exp, v + 2, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
}
// Callers run store_order() first, so the latch's writes are visible before
// the new version is.
__device__ __forceinline__ void ver_unlock(uint64_t *l) {
  __hip_atomic_fetch_add(l, (uint64_t)2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ver_unlock_obsolete(uint64_t *l) {
  __hip_atomic_fetch_add(l, (uint64_t)3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Packed key bytes as big-endian nibble words: 16 nibbles from nibble s of
// the key at p (8-byte aligned), first nibble in bits 63..60.  Reads only the
// aligned 8-byte words that hold nibbles below lim (the caller's bound on
// valid nibbles, s < lim), so every read stays inside the key's slot.
__device__ __forceinline__ uint64_t nib_word(const uint8_t *p, uint32_t s, uint32_t lim) {
  const uint32_t wi = s >> 4, sh = (s & 15u) * 4u;
  const uint64_t *w = (const uint64_t *)p;
  uint64_t hi = __builtin_bswap64(w[wi]);
  if (sh == 0) return hi;
  uint64_t lo = 16u * (wi + 1) < lim ? __builtin_bswap64(w[wi + 1]) : 0ull;
  return (hi << sh) | (lo >> (64u - sh));
}

struct Ctx {
  const uint8_t *kp;  // packed nibbles of the key (8-byte aligned)
  uint32_t len;       // hex length, terminator included
  uint32_t i;         // batch position
  uint32_t slot;      // rank among the batch's new keys (its node slots)
  bool fresh;         // the key does not exist yet
  uint32_t trie;
  uint64_t seq;
  uint64_t key_off;   // this key's packed bytes in the key arena
  uint64_t key_cap;   // key arena bytes in use (bounds for optimistic reads)
  uint32_t mark_from; // first node start position this lane marks (see mark())
  // the packed byte holding nibble j, loaded ahead of its use (0 when nibble
  // j is the terminator or past it)
  __device__ __forceinline__ uint32_t byte_at(uint32_t j) const { return j + 1 < len ? kp[j >> 1] : 0u; }
  __device__ __forceinline__ uint32_t nib(uint32_t j, uint32_t b) const {
    return j + 1 == len ? kNibbleT : (j & 1) ? (b & 15u) : (b >> 4);
  }
  __device__ __forceinline__ uint32_t nib(uint32_t j) const {
    return j + 1 == len ? kNibbleT : packed_nibble(kp, j);
  }
};

__device__ __forceinline__ void init_value(const NuArgs &a, const Ctx &c, uint32_t vid) {
  uint64_t *w = (uint64_t *)&a.P.value[vid];
  const int64_t vs = a.vidx[2 * c.i], ve = a.vidx[2 * c.i + 1];
  st(w + 0, a.vbase + (uint64_t)vs);                  // addr
  st(w + 1, a.hps ? (uint64_t)a.hps[c.i] : 0ull);     // hp
  st(w + 2, (uint64_t)(uint32_t)(ve - vs + 1));       // size, pad
  st(w + 3, c.seq);                                   // seq
}

// The existing value node `vid` takes this write's value (in place: the
// key's leaf is its only referrer, and the write is the key's only one in
// the batch).
__device__ __forceinline__ void rewrite_value(const NuArgs &a, const Ctx &c, uint32_t vid) {
  if (ld(&a.P.value[vid].seq) < c.seq) init_value(a, c, vid);
}

// A fresh short node, written as its eight words (layout of Short).
__device__ __forceinline__ void init_short(Short &S, uint64_t key_off, uint32_t key_start, uint32_t key_len,
                                           uint32_t depth, uint32_t term, ref_t child) {
  uint64_t *w = (uint64_t *)&S;
#pragma unroll
  for (int q = 0; q < 4; q++) st(w + q, 0ull);        // hash
  st(w + 4, key_off);
  st(w + 5, (uint64_t)child | (uint64_t)key_start << 32 | (uint64_t)key_len << 48);
  st(w + 6, (uint64_t)depth | (uint64_t)term << 24 | 1ull << 32);  // hash_size 0, dirty 1
  st(w + 7, 0ull);                                    // lock
}

// A fresh branch at depth d with two children, written as its sixteen words
// (layout of Branch).
__device__ __forceinline__ void init_branch(Branch &B, uint32_t d, uint32_t s0, ref_t r0, uint32_t s1, ref_t r1) {
  static_assert(offsetof(Branch, depth) == 34 && offsetof(Branch, lock) == 40 && offsetof(Branch, child) == 48,
                "init_branch writes Branch by words");
  uint64_t *w = (uint64_t *)&B;
#pragma unroll
  for (int q = 0; q < 4; q++) st(w + q, 0ull);        // hash
  st(w + 4, (uint64_t)d << 16 | 1ull << 32);          // hash_size 0, depth, dirty 1
  st(w + 5, 0ull);                                    // lock
#pragma unroll
  for (uint32_t q = 0; q < 8; q++) {                  // child[2q], child[2q + 1]
    const uint32_t lo = 2 * q, hi = 2 * q + 1;
    uint64_t x = (uint64_t)(lo == s0 ? r0 : lo == s1 ? r1 : 0u);
    x |= (uint64_t)(hi == s0 ? r0 : hi == s1 ? r1 : 0u) << 32;
    st(w + 6 + q, x);
  }
  st(w + 14, (uint64_t)(16 == s0 ? r0 : 16 == s1 ? r1 : 0u));  // child[16] (a key ending here), pad
  st(w + 15, 0ull);
}

// Publishes a node reference once the node's own stores have completed.
__device__ __forceinline__ void publish(ref_t *slot, ref_t r) {
  store_order();
  st(slot, r);
  store_order();
}

// New leaf for the key's nibbles [start, len): a short node over a value node,
// or the value node alone when nothing is left (key ends at a branch slot 16).
__device__ __forceinline__ ref_t make_leaf(const NuArgs &a, const Ctx &c, uint32_t start, bool *made_short) {
  uint32_t vid = a.base_value + c.slot;
  init_value(a, c, vid);
  ref_t v = make_ref(T_VALUE, vid);
  *made_short = start < c.len;
  if (!*made_short) return v;
  uint32_t sid = a.base_short + 2 * c.slot;
  init_short(a.P.shrt[sid], c.key_off, start, c.len - start, start, 1, v);
  return make_ref(T_SHORT, sid);
}

enum { ATT_DONE = 0, ATT_RESTART = 1 };

// A key classified as existing reached a point where it would create nodes
// (or a new key found itself present): the classification and the insert
// disagree.  Nothing is written; the insert reports an internal error.
__device__ __forceinline__ int misclassified(const NuArgs &a, const Ctx &c) {
  raise_error(a.P.cnt, ERR_INTERNAL, 21);
  return ATT_DONE;
}

// Dirty marks without atomics.  The keys are sorted and unique, so the lanes
// whose keys reach a node (every key with the node's nibble prefix) form one
// contiguous range of lanes; the first of them is the one lane whose key
// shares fewer nibbles with its predecessor's than the node's start position
// (the first key of a trie shares none).  That lane alone marks the node: a
// node flagged clean is flagged and staged for re-hashing.  Nodes this insert
// creates are born flagged and are staged by their creator.
__device__ __forceinline__ void mark(const Pools &P, const Ctx &c, DirtyStage &ds, ref_t node, uint32_t flag,
                                     uint32_t pos) {
  if (flag != 0 || pos < c.mark_from) return;
  __hip_atomic_store(node_dirty_flag(P, node), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  ds.push(P, node);
}

// Issues the loads of a node's version word (and a branch's dirty flag) for
// the next synchronisation point; other references get version 0.
__device__ __forceinline__ void head_issue(const Pools &P, ref_t r, uint64_t &v, uint32_t &flag) {
  const uint32_t ty = ref_type(r), idx = ref_idx(r);
  v = 0;
  flag = 1;
  if (ty == T_BRANCH && idx < P.cap_branch) {
    v = ld(&P.branch[idx].lock);
    flag = ld(&P.branch[idx].dirty);
  } else if (ty == T_SHORT && idx < P.cap_short) {
    v = ld(&P.shrt[idx].lock);
  }
}

// One optimistic attempt (optimistic lock coupling).  Returns ATT_DONE, or
// ATT_RESTART when a version check or latch upgrade failed (nothing was
// modified).  Each step costs two waits for memory: the node's contents, then
// the node's version re-check together with the child's version (the child
// reference was read under the node's version, so the re-check also covers
// it).  Nodes are never reused while the insert runs and a replaced node is
// left obsolete (bit 0), so a stale child reference always fails its own
// version check.  Path nodes are marked on the way down (a mark left by an
// abandoned attempt only costs a redundant re-hash).
__device__ __forceinline__ int attempt(const NuArgs &a, const Ctx &c, DirtyStage &ds) {
  const Pools &P = a.P;
  TrieHdr *H = &P.tries[c.trie];
  uint64_t *plock = &H->lock;
  uint64_t pver = ld(plock);
  load_order();
  if (pver & 3) return ATT_RESTART;
  ref_t *pslot = &H->root;
  ref_t parent = REF_NULL;
  ref_t node = ld(pslot);
  load_order();
  uint64_t v;
  uint32_t flag;
  uint32_t pos = 0, kb;  // kb: the key byte holding nibble pos
  {
    uint64_t hv = ld(plock);
    head_issue(P, node, v, flag);
    kb = c.byte_at(0);
    load_order();
    if (hv != pver || (v & 3)) return ATT_RESTART;
  }
  for (int guard = 0; guard < 2 * kMaxKeyNibbles + 4; guard++) {
    const uint32_t ty = ref_type(node), idx = ref_idx(node);
    if (node == REF_NULL) {
      if (!c.fresh) return misclassified(a, c);
      if (!ver_upgrade(plock, pver)) return ATT_RESTART;
      bool ms;
      ref_t leaf = make_leaf(a, c, pos, &ms);
      publish(pslot, leaf);
      ver_unlock(plock);
      if (ms) ds.push(P, leaf);
      return ATT_DONE;
    }
    if (ty == T_BRANCH) {
      if (idx >= P.cap_branch || pos >= c.len) return ATT_RESTART;
      Branch *B = &P.branch[idx];
      const uint32_t nib = c.nib(pos, kb);
      const ref_t child = ld(&B->child[nib]);
      load_order();
      uint64_t cv;
      uint32_t cflag;
      const uint64_t bv = ld(&B->lock);
      head_issue(P, child, cv, cflag);
      kb = c.byte_at(pos + 1);
      load_order();
      if (bv != v) return ATT_RESTART;
      mark(P, c, ds, node, flag, pos);
      if (cv & 3) return ATT_RESTART;
      parent = node;
      plock = &B->lock;
      pver = v;
      pslot = &B->child[nib];
      node = child;
      v = cv;
      flag = cflag;
      pos += 1;
      continue;
    }
    if (ty == T_SHORT) {
      if (idx >= P.cap_short) return ATT_RESTART;
      Short *S = &P.shrt[idx];
      // fields read as the words that hold them (offsets 32, 40, 48)
      const uint64_t koff = ld(&S->key_off);
      const uint64_t w40 = ld((const uint64_t *)((const uint8_t *)S + 40));
      const uint64_t w48 = ld((const uint64_t *)((const uint8_t *)S + 48));
      load_order();
      const ref_t child = (ref_t)(w40 & 0xffffffffu);
      const uint32_t kstart = (uint32_t)((w40 >> 32) & 0xffffu), klen = (uint32_t)(w40 >> 48);
      const uint32_t depth = (uint32_t)(w48 & 0xffffu), term = (uint32_t)((w48 >> 24) & 0xffu);
      const uint32_t sflag = (uint32_t)(w48 >> 32);
      if (klen == 0 || klen > kMaxKeyNibbles || depth != pos || term > 1 || pos >= c.len || koff > c.key_cap ||
          (uint64_t)(kstart + klen) / 2 + 1 > c.key_cap - koff)
        return ATT_RESTART;
      const uint8_t *sk = P.keys + (koff & ~7ull);
      // match the node's key against the key from pos, 16 nibbles per load
      const uint32_t sreal = klen - term, kleft = c.len - 1 - pos;
      const uint32_t mreal = min(sreal, kleft);
      uint32_t p = 0;
      while (p < mreal) {
        uint64_t x = nib_word(sk, kstart + p, kstart + sreal) ^ nib_word(c.kp, pos + p, c.len - 1);
        const uint32_t cnt = min(16u, mreal - p);
        if (cnt < 16) x &= ~0ull << (64u - 4u * cnt);
        if (x) {
          p += (uint32_t)__clzll((long long)x) >> 2;
          break;
        }
        p += cnt;
      }
      if (p == mreal && p == sreal && term && p == kleft) p++;  // both at the terminator
      if (p == klen) {
        if (ref_type(child) == T_VALUE) {
          // the key exists: its value node takes the new value in place,
          // with no latch (a split of S moves the reference, not the node)
          const uint32_t cv = ref_idx(child);
          const uint64_t sv = ld(&S->lock);
          load_order();
          if (sv != v) return ATT_RESTART;
          if (c.fresh || cv >= P.cap_value) return misclassified(a, c);
          rewrite_value(a, c, cv);
          mark(P, c, ds, node, sflag, pos);
          return ATT_DONE;
        }
        uint64_t cv;
        uint32_t cflag;
        const uint64_t sv = ld(&S->lock);
        head_issue(P, child, cv, cflag);
        kb = c.byte_at(pos + klen);
        load_order();
        if (sv != v) return ATT_RESTART;
        mark(P, c, ds, node, sflag, pos);
        if (cv & 3) return ATT_RESTART;
        parent = node;
        plock = &S->lock;
        pver = v;
        pslot = &S->child;
        node = child;
        v = cv;
        flag = cflag;
        pos += klen;
        continue;
      }
      if (pos + p >= c.len) return ATT_RESTART;
      if (!c.fresh) return misclassified(a, c);
      // split S at p: a new branch at depth pos + p (the latch upgrades
      // validate everything read above)
      const uint32_t d = pos + p;
      const uint32_t snib = (term && p + 1 == klen) ? kNibbleT : packed_nibble(sk, kstart + p);
      const uint32_t knib = c.nib(d);
      if (p == 0 && !ver_upgrade(plock, pver)) return ATT_RESTART;
      if (!ver_upgrade(&S->lock, v)) {
        if (p == 0) ver_unlock(plock);
        return ATT_RESTART;
      }
      const uint32_t bid = a.base_branch + c.slot;
      ref_t rest = child;
      if (p + 1 < klen) {
        const uint32_t rid = a.base_short + 2 * c.slot + 1;
        init_short(P.shrt[rid], koff, kstart + p + 1, klen - p - 1, d + 1, term, child);
        rest = make_ref(T_SHORT, rid);
        ds.push(P, rest);
      }
      bool ms;
      const ref_t leaf = make_leaf(a, c, d + 1, &ms);
      if (ms) ds.push(P, leaf);
      init_branch(P.branch[bid], d, snib, rest, knib, leaf);
      const ref_t bref = make_ref(T_BRANCH, bid);
      ds.push(P, bref);
      if (p == 0) {
        publish(pslot, bref);
        ver_unlock_obsolete(&S->lock);
        ver_unlock(plock);
      } else {
        // S keeps its first p nibbles and leads to the new branch: term
        // cleared (its own byte: the dirty word beside it may be written by
        // the lane that marks S), then child and key length in one word
        st(&S->term, (uint8_t)0);
        store_order();
        st((uint64_t *)((uint8_t *)S + 40), (uint64_t)bref | (uint64_t)kstart << 32 | (uint64_t)p << 48);
        store_order();
        ver_unlock(&S->lock);
        mark(P, c, ds, node, sflag, pos);
      }
      return ATT_DONE;
    }
    // a value node in a branch's slot 16: the key exists and takes the new
    // value in place (the reference to it was read under the branch's
    // version, validated with the node's own version check above)
    if (parent == REF_NULL || ref_type(parent) != T_BRANCH) return ATT_RESTART;
    const uint32_t vi = ref_idx(node);
    if (c.fresh || vi >= P.cap_value) return misclassified(a, c);
    rewrite_value(a, c, vi);
    return ATT_DONE;
  }
  return ATT_RESTART;
}

// Lane j inserts the j-th surviving write: batch position u_idx[j] (the last
// write of each distinct key, scal[SC_M_UNIQ] of them).
constexpr uint32_t kDirtyStage = 4096;  // staged dirty-list entries per block

// Existing keys.  The batch's unique keys are distinct, so a key exists
// during the insert exactly when it exists before it: a plain descent (no
// writer runs yet) sets fresh[j] = 1 for each key the insert must create.
__global__ __launch_bounds__(256) void k_lock_classify(NuArgs a, uint32_t *fresh) {
  const uint32_t j = grid_tid();
  if (j >= a.n) return;
  if (j >= a.scal[SC_M_UNIQ]) {
    fresh[j] = 0;
    return;
  }
  const Pools &P = a.P;
  const uint32_t i = a.u_idx[j];
  const uint64_t kr = a.kref[i];
  const uint8_t *kp = P.keys + a.kbase + (kr & ((1ull << 48) - 1));
  const uint32_t len = (uint32_t)(kr >> 48);
  const uint32_t t = a.tid ? (uint32_t)a.tid[i] : 0u;
  bool found = false;
  ref_t node = t < a.ntries ? P.tries[t].root : REF_NULL;
  uint32_t pos = 0;
  for (int guard = 0; guard < 2 * kMaxKeyNibbles + 4 && node != REF_NULL; guard++) {
    const uint32_t ty = ref_type(node), idx = ref_idx(node);
    if (ty == T_VALUE) {
      found = pos == len;
      break;
    }
    if (ty == T_BRANCH) {
      if (pos >= len || idx >= P.cap_branch) break;
      node = P.branch[idx].child[pos + 1 == len ? kNibbleT : packed_nibble(kp, pos)];
      pos++;
      continue;
    }
    if (idx >= P.cap_short) break;
    const Short &S = P.shrt[idx];
    const uint32_t klen = S.key_len, term = S.term;
    if (pos + klen > len) break;
    const uint32_t sreal = klen - term;
    // term: the node's key ends where the key does (a leaf of this key)
    if (term && pos + klen != len) break;
    if (!term && pos + klen == len) break;
    const uint8_t *sk = P.keys + (S.key_off & ~7ull);
    uint32_t p = 0;
    while (p < sreal) {
      uint64_t x = nib_word(sk, S.key_start + p, S.key_start + sreal) ^ nib_word(kp, pos + p, len - 1);
      const uint32_t cnt = min(16u, sreal - p);
      if (cnt < 16) x &= ~0ull << (64u - 4u * cnt);
      if (x) break;
      p += cnt;
    }
    if (p < sreal) break;
    pos += klen;
    node = S.child;
  }
  fresh[j] = found ? 0u : 1u;
}

__global__ __launch_bounds__(256, 5) void k_lock_insert(NuArgs a, const uint32_t *fresh, const uint32_t *rank,
                                                        uint64_t key_cap, uint32_t max_trips) {
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
  c.slot = active ? rank[j] : 0u;
  c.fresh = active && fresh[j] != 0;
  c.key_cap = key_cap;
  c.kp = a.P.keys;
  c.len = 1;
  c.trie = 0;
  c.seq = 0;
  c.key_off = 0;
  c.mark_from = 0;
  if (active) {
    const uint64_t kr = a.kref[i];
    c.key_off = a.kbase + (kr & ((1ull << 48) - 1));
    c.kp = a.P.keys + c.key_off;
    c.len = (uint32_t)(kr >> 48);
    c.trie = a.tid ? (uint32_t)a.tid[i] : 0u;
    c.seq = a.seq_base + i;
    c.mark_from = a.u_first[j] ? 0u : (uint32_t)a.u_lcp[j] + 1u;
  }
  bool done = !active;
  uint32_t restarts = 0;
  for (uint32_t trip = 0; trip < max_trips; trip++) {
    if (!done) {
      if (attempt(a, c, ds) == ATT_DONE)
        done = true;
      else
        restarts++;
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

// The slots the new keys took: per new key one branch, two short nodes and
// one value node.
__global__ void k_bump(Counters *cnt, const uint32_t *n_fresh) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const uint32_t k = *n_fresh;
    cnt->n_branch += k;
    cnt->n_short += 2 * k;
    cnt->n_value += k;
  }
}

}  // namespace

hipError_t lock_insert(const NuArgs &a, uint64_t key_cap, void *temp, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  // ecount / grp are free once the keys are sorted and deduplicated
  uint32_t *fresh = a.grp, *rank = a.ecount;
  uint32_t *n_fresh = a.scal + SC_NFRESH;
  k_lock_classify<<<nblk(a.n), kB, 0, s>>>(a, fresh);
  GMPT_LAUNCH_CHECK();
  hipError_t e = scan_exclusive_u32(fresh, rank, a.n, n_fresh, temp, s);
  if (e != hipSuccess) return e;
  k_lock_insert<<<nblk(a.n), kB, 0, s>>>(a, fresh, rank, key_cap, 1u << 20);
  GMPT_LAUNCH_CHECK();
  k_bump<<<1, 64, 0, s>>>(a.P.cnt, n_fresh);
  return hipGetLastError();
}

}  // namespace gmpt
