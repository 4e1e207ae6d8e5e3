// engine.hip -- host runtime of the forest: pools, staging, PhaseNU / PhaseHC.
#include "engine.h"

#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <new>

#include "get.h"
#include "locknu.h"
#include "phasehc.h"
#include "phasenu.h"
#include "prims.h"
#include "staging.h"

namespace gmpt {

namespace {
thread_local std::string g_err;

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Carves aligned sub-buffers out of one workspace allocation.
struct Carver {
  uint8_t *base;
  size_t off = 0;
  explicit Carver(void *b) : base((uint8_t *)b) {}
  template <typename T>
  T *take(size_t count) {
    off = align_up(off, 256);
    T *p = (T *)(base + off);
    off += count * sizeof(T);
    return p;
  }
};

size_t nu_ws_bytes(size_t n, bool tries) {
  size_t ne = 2 * n + 2;
  size_t b = 0;
  b += n * (4 + 8 + 8 + 16 + 8 + (tries ? 8 : 0) + 1 + 1 + 4 + 8 + 2 + sizeof(Stop) + 4 + sizeof(GroupInfo) + 4);
  b += ne * (4 + 16 + 4);  // leaf and branch records
  b += n * 8 * 5;      // sorted krefs; unique-order krefs and windows (two copies each)
  b += n * (4 * 7 + 1 + sizeof(Stop));
  b += ne * (4 + 4 + 4 + 4 + 8) + mt_bytes(align_up(ne, 256)) + 4 * (ne / 256 + 2);
  b += 2 * align_up(ne, 256) * (4 + 4 + 1) + 4 * (ne / 256 + 2);  // links between blocks of the build
  b += std::max(radix_temp_bytes(n), scan_temp_bytes(ne));
  b += 64 * 256;  // alignment slack for ~40 carves
  return b;
}

}  // namespace

void set_error(const std::string &e) { g_err = e; }
const char *last_error() { return g_err.c_str(); }

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e__ = (expr);                                                       \
    if (e__ != hipSuccess) {                                                       \
      set_error(std::string(#expr) + ": " + hipGetErrorString(e__));              \
      return -1;                                                                   \
    }                                                                              \
  } while (0)

Forest::Forest(int device, uint32_t ntries) : device_(device), ntries_(ntries) {
  if (hipSetDevice(device) != hipSuccess) {
    set_error("hipSetDevice failed");
    return;
  }
  if (hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) != hipSuccess) return;
  if (hipMalloc(&d_cnt_, sizeof(Counters)) != hipSuccess) return;
  if (hipHostMalloc(&h_cnt_, sizeof(Counters)) != hipSuccess) return;
  if (hipHostMalloc(&h_scal_, 64) != hipSuccess) return;
  if (hipMemset(d_cnt_, 0, sizeof(Counters)) != hipSuccess) return;
  if (hipEventCreate(&ev0_) != hipSuccess || hipEventCreate(&ev1_) != hipSuccess) return;
  if (hipEventCreate(&ev2_) != hipSuccess || hipEventCreate(&ev3_) != hipSuccess) return;
  if (hipEventCreateWithFlags(&ev_meta_, hipEventDisableTiming) != hipSuccess) return;
  if (hipEventCreateWithFlags(&tail_join_, hipEventDisableTiming) != hipSuccess) return;
  // default priority, like the main stream
  if (hipStreamCreateWithFlags(&side_.stream, hipStreamNonBlocking) != hipSuccess) return;
  if (hipEventCreateWithFlags(&side_.fork, hipEventDisableTiming) != hipSuccess) return;
  if (hipEventCreateWithFlags(&side_.join, hipEventDisableTiming) != hipSuccess) return;
  if (grow(tries_, sizeof(TrieHdr) * std::max<uint32_t>(ntries, 1), 0) != 0) return;
  if (hipMemset(tries_.p, 0, sizeof(TrieHdr) * std::max<uint32_t>(ntries, 1)) != hipSuccess) return;
  if (hipHostMalloc(&h_meta_, sizeof(HcMeta)) != hipSuccess) return;
  hc_spread_max_ = kHcSpreadMaxDefault;
  if (const char *e = getenv("GMPT_HC_SPREAD_MAX")) hc_spread_max_ = (uint32_t)strtoul(e, nullptr, 10);
  init_ok_ = true;
}

Forest::Forest(int device, uint32_t ntries, Arena *arena, HostMirror *host, hipStream_t stream)
    : device_(device), ntries_(ntries), arena_(arena), stream_(stream) {
  if (hipSetDevice(device) != hipSuccess) {
    set_error("hipSetDevice failed");
    return;
  }
  d_cnt_ = (Counters *)arena->take(sizeof(Counters));
  if (!d_cnt_) {
    set_error("device scratch region too small for the counters");
    return;
  }
  h_cnt_ = &host->cnt;
  h_scal_ = host->scal;
  h_meta_ = &host->meta;
  if (hipMemsetAsync(d_cnt_, 0, sizeof(Counters), stream_) != hipSuccess) return;
  if (grow(tries_, sizeof(TrieHdr) * std::max<uint32_t>(ntries, 1), 0) != 0) return;
  hc_spread_max_ = kHcSpreadMaxDefault;
  init_ok_ = true;
}

Forest::~Forest() {
  (void)hipSetDevice(device_);
  if (stream_ || arena_) (void)hipStreamSynchronize(stream_);
  if (arena_) return;  // nothing of its own to release
  drop_graph();
  release(plan_hist_);
  delete stager_;
  DevBuf *bufs[] = {&branch_, &shrt_, &value_, &keys_, &tries_, &dirty_, &st_keys_, &st_kidx_, &st_vidx_,
                    &st_hps_, &st_tid_, &ws_, &ws2_, &hc_sorted_, &hc_meta_, &hc_cursor_, &hc_slot_, &hc_enc_, &hc_scratch_,
                    &hc_lmeta_, &hc_lmeta_tmp_, &rw_};
  for (DevBuf *b : bufs) release(*b);
  for (DevBuf &b : vchunks_) release(b);
  if (d_cnt_) (void)hipFree(d_cnt_);
  if (h_cnt_) (void)hipHostFree(h_cnt_);
  if (h_scal_) (void)hipHostFree(h_scal_);
  if (h_meta_) (void)hipHostFree(h_meta_);
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  if (ev2_) (void)hipEventDestroy(ev2_);
  if (ev3_) (void)hipEventDestroy(ev3_);
  if (ev_meta_) (void)hipEventDestroy(ev_meta_);
  if (tail_join_) (void)hipEventDestroy(tail_join_);
  if (side_.fork) (void)hipEventDestroy(side_.fork);
  if (side_.join) (void)hipEventDestroy(side_.join);
  if (side_.stream) {
    (void)hipStreamSynchronize(side_.stream);
    (void)hipStreamDestroy(side_.stream);
  }
  if (stream_) (void)hipStreamDestroy(stream_);
}

int Forest::set_device() {
  root0_ok_ = false;  // any call but a root read may change the tries
  HIP_TRY(hipSetDevice(device_));
  return 0;
}

Pools Forest::pools() const {
  Pools P;
  P.branch = (Branch *)branch_.p;
  P.shrt = (Short *)shrt_.p;
  P.value = (Value *)value_.p;
  P.keys = (uint8_t *)keys_.p;
  P.tries = (TrieHdr *)tries_.p;
  P.dirty = (ref_t *)dirty_.p;
  P.cnt = d_cnt_;
  P.cap_branch = (uint32_t)std::min<size_t>(branch_.cap / sizeof(Branch), 0x3fffffffu);
  P.cap_short = (uint32_t)std::min<size_t>(shrt_.cap / sizeof(Short), 0x3fffffffu);
  P.cap_value = (uint32_t)std::min<size_t>(value_.cap / sizeof(Value), 0x3fffffffu);
  P.cap_dirty = (uint32_t)std::min<size_t>(dirty_.cap / sizeof(ref_t), 0xffffffffu);
  return P;
}

uint64_t Forest::pool_bytes_used() const {
  uint64_t v = 0;
  for (size_t i = 0; i < vchunks_.size(); i++) v += (i + 1 == vchunks_.size()) ? vchunk_used_ : vchunks_[i].cap;
  return (uint64_t)cnt_.n_branch * sizeof(Branch) + (uint64_t)cnt_.n_short * sizeof(Short) +
         (uint64_t)cnt_.n_value * sizeof(Value) + key_used_ + v;
}

uint64_t Forest::pool_bytes_reserved() const {
  uint64_t v = 0;
  for (const DevBuf &b : vchunks_) v += b.cap;
  return branch_.cap + shrt_.cap + value_.cap + keys_.cap + dirty_.cap + v;
}

bool Forest::is_pool(const DevBuf &b) const {
  return &b == &branch_ || &b == &shrt_ || &b == &value_ || &b == &keys_ || &b == &dirty_ || &b == &tries_;
}

void Forest::release(DevBuf &b) {
  if (b.p && !arena_) (void)(b.host ? hipHostFree(b.p) : hipFree(b.p));
  b.p = nullptr;
  b.cap = 0;
  b.host = false;
}

// Regrow a device buffer to at least `need` bytes, keeping the first `keep`.
// With pool_memory = host, the trie's pools grow in pinned host memory.
int Forest::grow(DevBuf &b, size_t need, size_t keep) {
  if (b.cap >= need) return 0;
  void *p = nullptr;
  size_t cap;
  const bool host = !arena_ && pool_host_ && is_pool(b);
  if (arena_) {
    // exactly what is asked: the region holds one call's forest
    cap = align_up(need, 256);
    if (!(p = arena_->take(cap))) return arena_exhausted();
  } else {
    cap = align_up(std::max(need, b.cap + b.cap / 2), 4096);
    hipError_t e = host ? hipHostMalloc(&p, cap, hipHostMallocDefault) : hipMalloc(&p, cap);
    if (e != hipSuccess) {
      set_error(std::string(host ? "host pool allocation failed: " : "device pool allocation failed: ") +
                hipGetErrorString(e));
      return -2;
    }
  }
  if (b.p && keep) HIP_TRY(hipMemcpyAsync(p, b.p, keep, hipMemcpyDefault, stream_));
  // fresh slots read as zero (empty children, unlocked, clean) even before a
  // writer publishes them, which keeps optimistic LockNU readers in bounds
  if (host) {
    HIP_TRY(hipStreamSynchronize(stream_));
    memset((uint8_t *)p + keep, 0, cap - keep);
  } else {
    HIP_TRY(hipMemsetAsync((uint8_t *)p + keep, 0, cap - keep, stream_));
  }
  HIP_TRY(hipStreamSynchronize(stream_));
  release(b);
  b.p = p;
  b.cap = cap;
  b.host = host;
  return 0;
}

int Forest::arena_exhausted() {
  set_error("the batch needs more device scratch than the fixed region holds (" + std::to_string(arena_->cap) +
            " bytes)");
  return -2;
}

int Forest::ensure_ws(DevBuf &b, size_t bytes) {
  if (b.cap >= bytes) return 0;
  if (arena_) {
    size_t cap = align_up(bytes, 256);
    if (!(b.p = arena_->take(cap))) return arena_exhausted();
    b.cap = cap;
    return 0;
  }
  if (b.p) {
    HIP_TRY(hipStreamSynchronize(stream_));
    release(b);
  }
  size_t cap = align_up(bytes + bytes / 4, 1 << 20);
  hipError_t e = hipMalloc(&b.p, cap);
  if (e != hipSuccess) {
    set_error(std::string("workspace allocation failed: ") + hipGetErrorString(e));
    return -2;
  }
  b.cap = cap;
  return 0;
}

int Forest::ensure_pools(uint64_t nb, uint64_t ns, uint64_t nv, uint64_t key_bytes) {
  int r;
  if ((r = grow(branch_, (cnt_.n_branch + nb) * sizeof(Branch), cnt_.n_branch * sizeof(Branch)))) return r;
  if ((r = grow(shrt_, (cnt_.n_short + ns) * sizeof(Short), cnt_.n_short * sizeof(Short)))) return r;
  if ((r = grow(value_, (cnt_.n_value + nv) * sizeof(Value), cnt_.n_value * sizeof(Value)))) return r;
  if ((r = grow(keys_, key_used_ + key_bytes + 64, key_used_))) return r;
  // every existing node is listed at most once, a build adds at most
  // 3 entries (holes included) per slot it reserved
  size_t dcap = ((size_t)cnt_.n_dirty + cnt_.n_branch + cnt_.n_short + 2 * nb + ns + 64) * sizeof(ref_t);
  if ((r = grow(dirty_, dcap, cnt_.n_dirty * sizeof(ref_t)))) return r;
  if ((uint64_t)(cnt_.n_branch + nb) > 0x3fffffffull || (uint64_t)(cnt_.n_short + ns) > 0x3fffffffull ||
      (uint64_t)(cnt_.n_value + nv) > 0x3fffffffull) {
    set_error("node pool exceeds 2^30 nodes of one kind");
    return -2;
  }
  return 0;
}

int Forest::sync_counters(hipEvent_t mark) {
  HIP_TRY(hipMemcpyAsync(h_cnt_, d_cnt_, sizeof(Counters), hipMemcpyDeviceToHost, stream_));
  if (mark) {
    // the host waits for the caller's event, recorded after the copy: one
    // packet instead of a synchronisation marker beside the event's own
    HIP_TRY(hipEventRecord(mark, stream_));
    HIP_TRY(hipEventSynchronize(mark));
  } else {
    HIP_TRY(hipStreamSynchronize(stream_));
  }
  cnt_ = *h_cnt_;
  if (cnt_.err != ERR_NONE) {
    uint32_t code = cnt_.err, arg = cnt_.err_arg;
    HIP_TRY(hipMemsetAsync(&d_cnt_->err, 0, 2 * sizeof(uint32_t), stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    const char *what = code == ERR_BAD_KEY ? "malformed key (needs hex nibbles < 16 and one terminator 16 at the end, "
                                             "at most 254 nibbles), trie id out of range, or rw flag other than 0 / 1"
                       : code == ERR_POOL  ? "node pool exhausted"
                                           : "internal error";
    set_error(std::string(what) + " (code " + std::to_string(code) + ", at " + std::to_string(arg) + ")");
    return code == ERR_BAD_KEY ? -3 : -4;
  }
  return 0;
}

uint8_t *Forest::value_space(uint64_t bytes) {
  bytes = align_up(bytes + 8, 256);
  if (arena_) {
    void *p = arena_->take(bytes);
    if (!p) arena_exhausted();
    return (uint8_t *)p;
  }
  if (vchunks_.empty() || vchunk_used_ + bytes > vchunks_.back().cap) {
    DevBuf c;
    c.cap = align_up(std::max<uint64_t>(bytes, 64ull << 20), 1 << 20);
    c.host = pool_host_;
    if ((c.host ? hipHostMalloc(&c.p, c.cap, hipHostMallocDefault) : hipMalloc(&c.p, c.cap)) != hipSuccess) {
      set_error("value arena allocation failed");
      return nullptr;
    }
    vchunks_.push_back(c);
    vchunk_used_ = 0;
  }
  uint8_t *p = (uint8_t *)vchunks_.back().p + vchunk_used_;
  vchunk_used_ += bytes;
  return p;
}

// Host batches are copied to device staging; values go into the value arena
// (device batches too, unless the caller lets the trie adopt its buffer).
int Forest::stage_batch(const BatchDesc &b, BatchDesc *d, uint64_t *vbase, bool copy_values, StageSeg *defer) {
  *d = b;
  uint32_t n = b.n;
  if (defer) *defer = StageSeg{nullptr, nullptr, 0};
  if (!b.on_device) {
    // extents of the host buffers: the largest end index of any pair (the
    // kernels reject pairs outside them)
    // and the key arena bytes the packed keys take (pairs may overlap, so
    // the key buffer's size does not bound it)
    uint64_t kb = 0, vb = 0, ks = 0;
    // large batches go through the pinned chunked stager (never in a forest
    // that may not allocate)
    const bool staged = !arena_ && 24ull * n >= HostStager::kMinBytes;
    if (staged && !stager_ && !(stager_ = new (std::nothrow) HostStager)) {
      set_error("out of host memory");
      return -2;
    }
    if (staged) {
      stager_->extents(b.keys_indexs, b.values_indexs, n, &kb, &vb, &ks, kMaxKeyNibbles);
    } else {
      for (uint32_t i = 0; i < n; i++) {
        const int64_t s = b.keys_indexs[2 * (size_t)i], e = b.keys_indexs[2 * (size_t)i + 1];
        if (e >= 0 && (uint64_t)e + 1 > kb) kb = (uint64_t)e + 1;
        const int64_t L = (s >= 0 && e >= s && e - s < kMaxKeyNibbles) ? e - s + 1 : 1;
        ks += (uint64_t)((L / 2 + 8) & ~7ll);
        if (b.values_indexs) {
          int64_t ve = b.values_indexs[2 * (size_t)i + 1];
          if (ve >= 0 && (uint64_t)ve + 1 > vb) vb = (uint64_t)ve + 1;
        }
      }
    }
    d->keys_bytes = kb;
    d->values_total = vb;
    d->key_slots = ks;
    int r;
    if ((r = ensure_ws(st_keys_, kb + 64))) return r;
    if ((r = ensure_ws(st_kidx_, 8ull * n + 64))) return r;
    if ((r = ensure_ws(st_vidx_, 16ull * n + 64))) return r;
    StageSeg seg[6];
    int ns = 0;
    seg[ns++] = {st_keys_.p, b.keys_hexs, kb};
    seg[ns++] = {st_kidx_.p, b.keys_indexs, 8ull * n};
    d->keys_hexs = (const uint8_t *)st_keys_.p;
    d->keys_indexs = (const int32_t *)st_kidx_.p;
    if (b.values_indexs) {
      seg[ns++] = {st_vidx_.p, b.values_indexs, 16ull * n};
      d->values_indexs = (const int64_t *)st_vidx_.p;
    }
    if (b.values_hps) {
      if ((r = ensure_ws(st_hps_, 8ull * n + 64))) return r;
      seg[ns++] = {st_hps_.p, b.values_hps, 8ull * n};
      d->values_hps = (const uint64_t *)st_hps_.p;
    }
    if (b.trie_ids) {
      if ((r = ensure_ws(st_tid_, 4ull * n + 64))) return r;
      seg[ns++] = {st_tid_.p, b.trie_ids, 4ull * n};
      d->trie_ids = (const int32_t *)st_tid_.p;
    }
    d->on_device = true;
    if (copy_values) {
      uint8_t *vs = value_space(vb);
      if (!vs) return -2;
      if (staged && defer)
        *defer = StageSeg{vs, b.values_bytes, vb};
      else
        seg[ns++] = {vs, b.values_bytes, vb};
      d->values_bytes = vs;
    }
    if (staged) {
      HIP_TRY(stager_->copy(stream_, seg, ns));
    } else {
      for (int q = 0; q < ns; q++)
        if (seg[q].bytes) HIP_TRY(hipMemcpyAsync(seg[q].dst, seg[q].src, seg[q].bytes, hipMemcpyDefault, stream_));
    }
  } else if (copy_values && !b.adopt_values) {
    uint8_t *vs = value_space(b.values_total);
    if (!vs) return -2;
    if (b.values_total)
      HIP_TRY(hipMemcpyAsync(vs, b.values_bytes, b.values_total, hipMemcpyDefault, stream_));
    d->values_bytes = vs;
  }
  // device batches: the packed keys of non-overlapping pairs; the packing
  // pass reports a larger total (kErrArgKeyArena) and the insert is re-run
  if (b.on_device) d->key_slots = b.keys_bytes / 2 + 8ull * n;
  *vbase = (uint64_t)(uintptr_t)d->values_bytes;
  return 0;
}

// After a failed synchronisation: was it the packing pass finding the key
// arena too small?  Then the arena is grown to the batch's packed total
// (kept in the scalars) and true is returned, so the caller runs the insert
// again.  The failed run changed nothing but dirty flags of path nodes.
bool Forest::grow_key_arena_after_error() {
  if (cnt_.err != ERR_POOL || cnt_.err_arg != kErrArgKeyArena) return false;
  const uint64_t ktot = *(const uint64_t *)(cnt_.scal + SC_KTOT);
  return grow(keys_, key_used_ + ktot + 64, key_used_) == 0;
}

// Workspace of one PhaseNU-style call: every per-key array of NuArgs carved
// from ws_ (sized by nu_ws_bytes), batch pointers and slot bases filled in.
int Forest::carve_nu(NuArgs &a, const BatchDesc &d, uint64_t vbase, bool has_tid, void **temp) {
  const uint32_t n = d.n;
  int r;
  rec_ = LeafRecords();  // the workspace is about to be reused and the trie changed
  if ((r = ensure_ws(ws_, nu_ws_bytes(n, has_tid)))) return r;
  Carver c(ws_.p);
  memset(&a, 0, sizeof(a));
  a.P = pools();
  a.kh = d.keys_hexs;
  a.kidx = d.keys_indexs;
  a.vidx = d.values_indexs;
  a.hps = d.values_hps;
  a.tid = has_tid ? d.trie_ids : nullptr;
  a.n = n;
  a.ntries = ntries_;
  a.kh_bytes = d.keys_bytes;
  a.v_bytes = d.values_total;
  a.kbase = key_used_;
  a.key_cap = keys_.cap;
  a.vbase = vbase;
  a.base_branch = cnt_.n_branch;
  a.base_short = cnt_.n_short;
  a.base_value = cnt_.n_value;
  size_t ne = 2 * (size_t)n + 2;
  a.klen = c.take<uint32_t>(n);
  a.kpoff = c.take<uint64_t>(n);
  a.kref = c.take<uint64_t>(n);
  a.sk0 = c.take<uint64_t>(n);
  a.sk1 = c.take<uint64_t>(n);
  a.sv0 = c.take<uint32_t>(n);
  a.sv1 = c.take<uint32_t>(n);
  if (has_tid) {
    a.tk0 = c.take<uint32_t>(n);
    a.tk1 = c.take<uint32_t>(n);
  }
  a.scal = d_cnt_->scal;  // copied back with the counters
  a.lcp = c.take<uint8_t>(n);
  a.flags = c.take<uint8_t>(n);
  a.run = c.take<uint32_t>(n);
  a.u_idx = c.take<uint32_t>(n);
  a.u_trie = c.take<uint32_t>(n);
  a.u_lcp = c.take<uint8_t>(n);
  a.u_first = c.take<uint8_t>(n);
  a.stop = c.take<Stop>(n);
  a.grp = c.take<uint32_t>(n);
  a.ginfo = c.take<GroupInfo>(n);
  a.ecount = c.take<uint32_t>(n);
  a.merge = c.take<uint32_t>(n);
  a.u_idx2 = c.take<uint32_t>(n);
  a.skref = c.take<uint64_t>(n);
  a.u_kref = c.take<uint64_t>(n);
  a.u_kref2 = c.take<uint64_t>(n);
  a.u_win = c.take<uint64_t>(n);
  a.u_win2 = c.take<uint64_t>(n);
  a.u_trie2 = c.take<uint32_t>(n);
  a.perm0 = c.take<uint32_t>(n);
  a.perm1 = c.take<uint32_t>(n);
  a.gkey0 = c.take<uint32_t>(n);
  a.gkey1 = c.take<uint32_t>(n);
  a.u_lcp2 = c.take<uint8_t>(n);
  a.stop2 = c.take<Stop>(n);
  a.e_src = c.take<uint32_t>(ne);
  a.e_grp = c.take<uint32_t>(ne);
  a.leaf_pk = c.take<uint32_t>(ne);
  a.leaf_meta = c.take<uint64_t>(2 * ne);
  a.bdesc = c.take<uint32_t>(ne);
  a.st_stride = align_up(ne, 256);
  a.st_levels = mt_levels(a.st_stride);
  a.st = c.take<uint8_t>(mt_bytes(a.st_stride));
  a.repof = c.take<uint32_t>(ne);
  a.bid = c.take<uint32_t>(ne);
  a.plohi = c.take<uint64_t>(ne);
  a.bcnt = c.take<uint32_t>(ne / 256 + 2);
  a.xp = c.take<uint32_t>(2 * align_up(ne, 256));
  a.xm = c.take<uint32_t>(2 * align_up(ne, 256));
  a.xs = c.take<uint8_t>(2 * align_up(ne, 256));
  a.xcnt = c.take<uint32_t>(ne / 256 + 2);
  *temp = c.take<uint8_t>(std::max(radix_temp_bytes(n), scan_temp_bytes(ne)));
  if (c.off > ws_.cap) {
    set_error("workspace carve overflow");
    return -4;
  }
  HIP_TRY(nu_init_scalars(a, stream_));
  return 0;
}

// Sorts the packed keys (by trie, then nibble path) and keeps the last
// occurrence of each key: u_idx[0 .. scal[SC_M_UNIQ]) in sorted order.  pass 0
// sorts by the whole windows (top-8-bit buckets sorted in LDS; above
// kBucketSortMaxN keys, by the windows' leading bits with a fix-up of equal
// runs), pass 1 by every window bit that varies across the batch, pass 2 by
// the whole keys.
// Passes 0 and 2 need no host synchronisation: the caller reads
// flag_fallback at its next one and, if a flag is set, runs next_pass() and
// repeats the work that followed.
int Forest::sort_unique(NuArgs &a, uint32_t trie_bits, void *temp, int pass, int *cur) {
  uint64_t *sk = nullptr;
  uint32_t *sv = nullptr;
  int nwords = 0;
  int r;
  if (pass == 2) {
    if ((r = sync_counters())) return r;
    uint32_t W = cnt_.scal[SC_W], maxlen = cnt_.scal[SC_MAXLEN];
    nwords = (int)((maxlen > W ? maxlen - W : 1) + 15) / 16;
  }
  uint64_t vmask = 0, sk_mask = 0;
  if (pass == 1 && 4ull * ((a.n + 255) / 256) <= a.n) {
    // keys with long shared stretches inside the window (ASCII digits, fixed
    // fields) sort by fewer radix digits once the constant bits are dropped
    HIP_TRY(nu_window_mask(a, *cur ? a.sk1 : a.sk0, stream_));
    if ((r = sync_counters())) return r;
    const uint64_t wor = *(const uint64_t *)(cnt_.scal + SC_WOR), wand = *(const uint64_t *)(cnt_.scal + SC_WAND);
    const uint64_t m = wor & ~wand;
    const int pop = __builtin_popcountll(m);
    if (pop >= 1 && pop <= 56) vmask = m;  // saves at least one 8-bit digit
  }
  HIP_TRY(nu_sort(a, nwords, trie_bits, temp, stream_, &sk, &sv, pass == 1, cur, vmask, &sk_mask));
  // the sorted keys are the plain windows unless the whole keys were sorted
  // or the windows were compressed to their varying bits
  a.u_win_ok = pass != 2 && sk_mask == 0;
  if (pass > 0) HIP_TRY(nu_reset_dup(a, stream_));
  HIP_TRY(nu_adjacent(a, sk, sv, pass == 2 ? 1 : 0, stream_, sk_mask));
  HIP_TRY(nu_unique(a, sk, sv, temp, stream_));
  return 0;
}

// After a synchronisation: the sort pass to redo with, or -1 when the last
// sort stands.  A run of equal leading window bits too long to fix up asks
// for a sort by every varying window bit; two distinct keys with equal
// windows ask for a sort by whole keys.
int Forest::next_pass(int pass) {
  uint32_t f = cnt_.flag_fallback;
  if (f == 0 || pass == 2) return -1;
  if (f & 1u) {
    stats_.fallback_sorts++;
    return 2;
  }
  return 1;
}

int Forest::puts_2phase(const BatchDesc &b, void **targets, int *n_targets) {
  int r;
  if ((r = set_device())) return r;
  if (b.n > 0x3fffffffu) {
    set_error("batch too large");
    return -3;
  }
  if (b.on_device && b.n && (b.keys_bytes == 0)) {
    set_error("device batch needs keys_bytes and values_total");
    return -3;
  }
  stats_.fallback_sorts = 0;
  stats_.regroups = 0;
  if (b.n > 0) {
    BatchDesc d;
    uint64_t vbase = 0;
    StageSeg vals{nullptr, nullptr, 0};  // a large host batch's values, staged while the insert kernels run
    if ((r = stage_batch(b, &d, &vbase, true, side_.stream ? &vals : nullptr))) return r;
    // the values' copies go on the second stream, after all work queued so
    // far on the forest's stream (not after the insert kernels)
    if (vals.bytes) {
      HIP_TRY(hipEventRecord(side_.fork, stream_));
      HIP_TRY(hipStreamWaitEvent(side_.stream, side_.fork, 0));
    }
    const uint32_t n = b.n;
    // slots: branches <= 2n+2, shorts <= 2(2n+2), values: 2n+2 for elements + n for updates
    if ((r = ensure_pools(2ull * n + 2, 4ull * n + 4, 3ull * n + 2, d.key_slots))) return r;
    const bool has_tid = d.trie_ids != nullptr && ntries_ > 1;
    // a bulk build into a forest of empty tries skips the descent and grouping
    const bool empty = all_empty_;
    all_empty_ = false;
    NuArgs a;
    void *temp = nullptr;
    uint32_t trie_bits = 0;
    if (has_tid) {
      while ((1ull << trie_bits) < ntries_) trie_bits++;
    }
    for (int run = 0;; run++) {
      if ((r = carve_nu(a, d, vbase, has_tid, &temp))) return r;
      a.e_ident = empty && !has_tid ? 1u : 0u;
      HIP_TRY(nu_prepare_keys(a, temp, stream_));
      int cur = 0;
      for (int pass = 0; pass >= 0;) {
        if ((r = sort_unique(a, trie_bits, temp, pass, &cur))) break;
        HIP_TRY(nu_descend_and_group(a, temp, stream_, empty));
        if (empty) {
          // into empty tries no group needs merging: the build follows at
          // once (its kernels skip themselves while the sort asks for another
          // pass or an error is pending), one synchronisation per pass
          HIP_TRY(nu_emit(a, temp, stream_, true));
          HIP_TRY(nu_build(a, temp, stream_, true));
        }
        if (vals.bytes) {
          // no insert kernel reads value bytes (only their index pairs): the
          // values cross PCIe on the second stream while they run, and the
          // forest's stream waits for them before anything queued later
          HIP_TRY(stager_->copy(side_.stream, &vals, 1));
          HIP_TRY(hipEventRecord(side_.join, side_.stream));
          HIP_TRY(hipStreamWaitEvent(stream_, side_.join, 0));
          vals.bytes = 0;
        }
        if ((r = sync_counters())) break;
        pass = next_pass(pass);
        if (pass >= 0) {
          // redo from the sort; marks the discarded descent made are on the
          // batch's own paths and stay valid
          HIP_TRY(hipMemsetAsync(&d_cnt_->flag_fallback, 0, 4, stream_));
          HIP_TRY(hipMemsetAsync(a.scal + SC_MERGE, 0, 4, stream_));
        }
      }
      if (r == 0) break;
      if (run > 0 || !grow_key_arena_after_error()) return r;
    }
    if (!empty) {
      if (cnt_.scal[SC_MERGE]) {
        // a split group straddles a subtree: make every group contiguous
        uint32_t m = cnt_.scal[SC_M_UNIQ];
        HIP_TRY(nu_regroup(a, m, m, temp, stream_));
        stats_.regroups++;
      }
      HIP_TRY(nu_emit(a, temp, stream_, false));
      HIP_TRY(nu_build(a, temp, stream_, false));
      if ((r = sync_counters())) return r;
    }
    // the batch's packed keys stay only if a new node refers to them: a batch
    // of updates alone (no elements) leaves them unused, so the arena keeps
    // its place (keys of repeated updates do not accumulate)
    if (cnt_.scal[SC_NE] != 0) {
      key_used_ += *(const uint64_t *)(cnt_.scal + SC_KTOT);
      key_used_ = align_up(key_used_, 8);
    }
    // the new leaves' hashing bins, valid until the next change to the trie
    rec_.pk = a.leaf_pk;
    rec_.meta = a.leaf_meta;
    rec_.lo = cnt_.scal[SC_DBASE];
    rec_.n = cnt_.scal[SC_NE];
    rec_.bdesc = a.bdesc;
    rec_.blo = rec_.lo + rec_.n;
    rec_.bn = cnt_.scal[SC_NREP_TOT];
  }
  if (targets) *targets = dirty_.p;
  if (n_targets) *n_targets = (int)cnt_.n_dirty;
  return 0;
}

int Forest::hash(void *targets, int n_targets) {
  int r;
  if ((r = set_device())) return r;
  if (targets && targets != dirty_.p) {
    set_error("hash targets do not belong to this trie");
    return -3;
  }
  (void)n_targets;  // the whole dirty list is hashed: a subset would leave stale ancestors
  uint32_t n = cnt_.n_dirty;
  stats_.hash_launches = 0;
  stats_.hashed_nodes = 0;
  stats_.leaf_nodes = 0;
  stats_.ext_direct_nodes = 0;
  stats_.last_hash_ms = 0;
  unsigned long long *parts = nullptr;  // the pass's work counts (hc_count), folded by hc_roots
  bool leaf_from_ev0 = false;           // the leaf launch's start event is ev0_
  if (n > 0) {
    Pools P = pools();
    if ((r = ensure_ws(hc_sorted_, sizeof(ref_t) * (size_t)n))) return r;
    if ((r = ensure_ws(hc_slot_, sizeof(uint64_t) * (size_t)n))) return r;
    if ((r = ensure_ws(hc_enc_, sizeof(uint32_t) * (size_t)n))) return r;
    if ((r = ensure_ws(hc_meta_, sizeof(HcMeta) + 32 * kHcCountParts))) return r;
    if ((r = ensure_ws(hc_cursor_, sizeof(uint32_t) * kHcBins))) return r;
    if ((r = ensure_ws(ws2_, hc_temp_bytes(n)))) return r;
    ref_t *sorted = (ref_t *)hc_sorted_.p;
    uint64_t *slot = (uint64_t *)hc_slot_.p;
    uint32_t *encl = (uint32_t *)hc_enc_.p;
    uint64_t *lmeta = nullptr, *lmeta_tmp = nullptr;
    if (hc_fuse_leaves_) {
      if ((r = ensure_ws(hc_lmeta_, 16 * (size_t)n))) return r;
      if ((r = ensure_ws(hc_lmeta_tmp_, 16 * (size_t)n))) return r;
      lmeta = (uint64_t *)hc_lmeta_.p;
      lmeta_tmp = (uint64_t *)hc_lmeta_tmp_.p;
    }
    HIP_TRY(hc_classify(P, (const ref_t *)dirty_.p, n, (HcMeta *)hc_meta_.p, (uint32_t *)hc_cursor_.p, sorted, slot,
                        encl, ws2_.p, stream_, lmeta_tmp, lmeta, rec_));
    HIP_TRY(hipMemcpyAsync(h_meta_, hc_meta_.p, sizeof(HcMeta), hipMemcpyDeviceToHost, stream_));
    // One event marks the bins' copy (the host waits for it), the start of the
    // hash timing and, with the direct leaf kernel, the start of that launch:
    // each record is a packet the GPU processes between two kernels.
    meta_mark_ = ev0_ ? ev0_ : ev_meta_;
    if (meta_mark_) HIP_TRY(hipEventRecord(meta_mark_, stream_));
    if (lmeta) {
      // the leaves of level 0 (its bins): launched before the bins reach the
      // host, so the GPU hashes them while the host plans the other levels
      leaf_from_ev0 = true;
      // the launch's bound on the leaves: the last build's branch and
      // extension entries, when they end the list, hold none
      const uint32_t n_upper = rec_.bdesc && rec_.blo + 2ull * rec_.bn == n ? n - 2 * rec_.bn : n;
      HIP_TRY(hc_hash_leaves_direct(P, sorted, lmeta, (const HcMeta *)hc_meta_.p, n_upper, encl, stream_));
      if (ev3_) HIP_TRY(hipEventRecord(ev3_, stream_));
    }
    if (meta_mark_)
      HIP_TRY(hipEventSynchronize(meta_mark_));
    else
      HIP_TRY(hipStreamSynchronize(stream_));
    const HcMeta &M = *(const HcMeta *)h_meta_;
    if ((r = ensure_ws(hc_scratch_, M.scratch_bytes + 64))) return r;
    uint32_t n_leaf = 0;
    if ((r = hash_levels(P, M.hist, lmeta != nullptr, false, &n_leaf))) return r;
    parts = (unsigned long long *)((uint8_t *)hc_meta_.p + sizeof(HcMeta));
    HIP_TRY(hc_count(encl, (const HcMeta *)hc_meta_.p, n, n_leaf, parts, stream_));
    stats_.leaf_nodes = n_leaf;
  }
  HIP_TRY(hc_roots(pools(), ntries_, parts, stream_));  // also empties the dirty list
  rec_ = LeafRecords();
  // ev1_ closes the hash timing after the roots and the counters' read
  if ((r = sync_counters(n > 0 && ev0_ ? ev1_ : nullptr))) return r;
  // the counters are zeroed by the classification, so only a pass that ran sets them
  stats_.perms = n ? cnt_.n_perms : 0;
  stats_.enc_bytes = n ? cnt_.enc_bytes : 0;
  stats_.leaf_perms = n ? cnt_.leaf_perms : 0;
  stats_.leaf_bytes = n ? cnt_.leaf_bytes : 0;
  stats_.leaf_perm_ms = 0;
  if (n > 0 && ev0_) {
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ev0_, ev1_));
    stats_.last_hash_ms = ms;
    if (stats_.leaf_nodes) {
      HIP_TRY(hipEventElapsedTime(&ms, leaf_from_ev0 ? ev0_ : ev2_, ev3_));
      stats_.leaf_perm_ms = ms;
    }
  }
  root0_ok_ = cnt_.root0_size <= 32;
  return 0;
}

// The hash levels after the leaves the direct kernel takes, from the bins
// (hist, kHcBins counts), one range per non-empty level; captured = true for
// a pass being captured into a graph (no timing events, no second stream).
int Forest::hash_levels(const Pools &P, const uint32_t *hist, bool direct_leaves, bool captured,
                        uint32_t *n_leaf_out) {
  ref_t *sorted = (ref_t *)hc_sorted_.p;
  uint64_t *slot = (uint64_t *)hc_slot_.p;
  uint32_t *encl = (uint32_t *)hc_enc_.p;
  uint64_t *scratch = (uint64_t *)hc_scratch_.p;
  uint32_t n_direct = 0;  // the leaves of level 0, all hashed by the direct kernel
  if (direct_leaves)
    for (uint32_t c = 0; c < kHcClasses; c++) n_direct += hist[c];
  const auto level_count = [&](uint32_t lv) {
    uint32_t cnt = 0;
    for (uint32_t c = 0; c < kHcClasses; c++) cnt += hist[lv * kHcClasses + c];
    return cnt;
  };
  // The tail: the deepest inner levels (levels 2 .. tail_end - 1), when they
  // are small.  They depend on nothing but their own leaf children and each
  // other, so the second stream hashes those leaves and then the tail levels
  // while the leaf-level launch runs; the forest's stream waits for them
  // before the first level above the tail.  Leaves the word builder cannot
  // take (level 1) are hashed after the leaf launch, so a pass with any
  // hashes no tail.
  uint32_t tail_end = 2, tail_n = 0;
  if (direct_leaves && n_direct && !captured && side_.stream && meta_mark_ && tail_join_ && level_count(1) == 0) {
    for (uint32_t lv = 2; lv < kHcLevels; lv++) {
      const uint32_t cnt = level_count(lv);
      if (cnt > kTailLevelMax || tail_n + cnt > kTailMax) break;
      tail_n += cnt;
      tail_end = lv + 1;
    }
  }
  const bool tail = tail_n > 0;
  bool joined = !tail;
  uint32_t start = 0, n_leaf = 0;
  for (uint32_t lv = 0; lv < kHcLevels; lv++) {
    const uint32_t cnt = level_count(lv);
    const bool in_tail = tail && lv >= 2 && lv < tail_end;
    if (tail && lv == 2) {
      // after the classification (its bins' copy), beside the leaf launch
      HIP_TRY(hipStreamWaitEvent(side_.stream, meta_mark_, 0));
      HIP_TRY(hc_tail_leaves(P, sorted + start, tail_n, side_.stream));
    }
    if (cnt && !in_tail && !joined && lv >= 2) {
      HIP_TRY(hipStreamWaitEvent(stream_, tail_join_, 0));
      joined = true;
    }
    if (cnt) {
      const bool leaves = lv <= 1;  // level 1: leaves the direct kernel does not take
      const uint32_t skip = lv == 0 ? n_direct : 0u;  // already queued
      const bool timed = lv == 0 && !direct_leaves && ev2_ && !captured;
      const uint32_t n_ext = leaves ? 0u : hist[lv * kHcClasses + kHcClasses - 1];
      if (cnt > skip)
        HIP_TRY(hc_hash_range(P, sorted + start + skip, slot + start + skip, cnt - skip, scratch,
                              encl + start + skip, leaves, hc_spread_max_, in_tail ? side_.stream : stream_,
                              timed ? ev2_ : nullptr, timed ? ev3_ : nullptr, n_ext,
                              hc_fuse_inner_ ? &hist[lv * kHcClasses] : nullptr));
      stats_.hash_launches++;
      stats_.hashed_nodes += cnt;
      if (!leaves && n_ext >= kExtDirectMin && cnt > hc_spread_max_) stats_.ext_direct_nodes += n_ext;
      if (leaves) n_leaf += cnt;
    }
    if (tail && lv + 1 == tail_end) HIP_TRY(hipEventRecord(tail_join_, side_.stream));
    start += cnt;
  }
  if (!joined) HIP_TRY(hipStreamWaitEvent(stream_, tail_join_, 0));  // every inner level was in the tail
  *n_leaf_out = n_leaf;
  return 0;
}

int Forest::root_hash(uint32_t trie, uint8_t *out, int *size) {
  if (trie == 0 && ntries_ > 0 && root0_ok_) {  // no call since the last hash pass, whose counter read brought it
    *size = (int)cnt_.root0_size;
    if (out) memcpy(out, cnt_.root0, cnt_.root0_size);
    return 0;
  }
  int r;
  if ((r = set_device())) return r;
  if (trie >= ntries_) {
    set_error("trie id out of range");
    return -3;
  }
  TrieHdr h;
  HIP_TRY(hipMemcpyAsync(&h, (TrieHdr *)tries_.p + trie, sizeof(TrieHdr), hipMemcpyDefault, stream_));
  HIP_TRY(hipStreamSynchronize(stream_));
  *size = (int)h.root_hash_size;
  if (out) memcpy(out, h.root_hash, h.root_hash_size);
  return 0;
}

int Forest::root_hashes(uint32_t first, uint32_t count, uint8_t *out, int32_t *sizes) {
  int r;
  if ((r = set_device())) return r;
  if (first > ntries_ || count > ntries_ - first) {
    set_error("trie range out of range");
    return -3;
  }
  if (count == 0) return 0;
  std::vector<TrieHdr> h(count);
  HIP_TRY(hipMemcpyAsync(h.data(), (TrieHdr *)tries_.p + first, sizeof(TrieHdr) * count, hipMemcpyDefault,
                         stream_));
  HIP_TRY(hipStreamSynchronize(stream_));
  for (uint32_t i = 0; i < count; i++) {
    sizes[i] = (int32_t)h[i].root_hash_size;
    memcpy(out + 32 * (size_t)i, h[i].root_hash, 32);
  }
  return 0;
}

int Forest::reset_trie(uint32_t trie) {
  int r;
  if ((r = set_device())) return r;
  if (trie >= ntries_) {
    set_error("trie id out of range");
    return -3;
  }
  rec_ = LeafRecords();
  HIP_TRY(hipMemsetAsync((TrieHdr *)tries_.p + trie, 0, sizeof(TrieHdr), stream_));
  HIP_TRY(hipStreamSynchronize(stream_));
  return 0;
}

int Forest::clear() {
  int r;
  if ((r = set_device())) return r;
  rec_ = LeafRecords();
  HIP_TRY(hc_clear(d_cnt_, (TrieHdr *)tries_.p, ntries_, stream_));
  memset(&cnt_, 0, sizeof(cnt_));
  all_empty_ = true;
  key_used_ = 0;
  // keep the first value chunk, release the others
  for (size_t i = 1; i < vchunks_.size(); i++) release(vchunks_[i]);
  if (vchunks_.size() > 1) vchunks_.resize(1);
  vchunk_used_ = 0;
  return 0;
}

int Forest::set_option(const std::string &name, int64_t value) {
  if (name == "hc_spread_max") {
    if (value < 0 || value > 0xffffffffll) {
      set_error("hc_spread_max out of range");
      return -3;
    }
    hc_spread_max_ = (uint32_t)value;
    return 0;
  }
  if (name == "pool_memory") {
    if (value != 0 && value != 1) {
      set_error("pool_memory must be 0 (device) or 1 (host)");
      return -3;
    }
    if (arena_ && value) {
      set_error("a forest over a fixed device region keeps its pools there");
      return -3;
    }
    pool_host_ = value != 0;
    return 0;
  }
  if (name == "hc_fuse_leaves" || name == "hc_fuse_inner") {
    if (value != 0 && value != 1) {
      set_error(name + " must be 0 or 1");
      return -3;
    }
    (name == "hc_fuse_leaves" ? hc_fuse_leaves_ : hc_fuse_inner_) = value != 0;
    return 0;
  }
  set_error("unknown option: " + name);
  return -3;
}

int Forest::get_option(const std::string &name, int64_t *value) const {
  if (name == "hc_spread_max") {
    *value = hc_spread_max_;
    return 0;
  }
  if (name == "hc_fuse_leaves" || name == "hc_fuse_inner") {
    *value = (name == "hc_fuse_leaves" ? hc_fuse_leaves_ : hc_fuse_inner_) ? 1 : 0;
    return 0;
  }
  if (name == "pool_memory") {
    *value = pool_host_ ? 1 : 0;
    return 0;
  }
  set_error("unknown option: " + name);
  return -3;
}

int Forest::autotune(const BatchDesc &sample, int mode, int64_t *chosen, double *best_ms) {
  static const uint32_t kCand[] = {0, 256, 1024, 4096, 16384, 65536};
  Forest t(device_, ntries_);
  if (!t.ok()) return -1;
  double best = 1e30;
  uint32_t arg = hc_spread_max_;
  for (uint32_t c : kCand) {
    t.hc_spread_max_ = c;
    t.hc_fuse_leaves_ = hc_fuse_leaves_;
    t.hc_fuse_inner_ = hc_fuse_inner_;
    double ms = 1e30;
    for (int rep = 0; rep < 4; rep++) {  // the first run warms the pools; the best of the other three counts
      int r;
      void *tg = nullptr;
      int nt = 0;
      if ((r = t.clear())) return r;
      r = mode == 0 ? t.puts_2phase(sample, &tg, &nt) : t.puts_latching(sample, &tg, &nt);
      if (r) return r;
      if ((r = t.hash(tg, nt))) return r;
      if (rep > 0) ms = std::min(ms, (double)t.stats().last_hash_ms);
    }
    if (ms < best) {
      best = ms;
      arg = c;
    }
  }
  hc_spread_max_ = arg;
  if (chosen) *chosen = arg;
  if (best_ms) *best_ms = best;
  return 0;
}

int Forest::synchronize() {
  int r;
  if ((r = set_device())) return r;
  HIP_TRY(hipStreamSynchronize(stream_));
  return 0;
}

int Forest::gets(const BatchDesc &keys, uint64_t *hps, int32_t *sizes) {
  int r;
  if ((r = set_device())) return r;
  uint32_t n = keys.n;
  if (n == 0) return 0;
  BatchDesc d;
  uint64_t vb;
  if ((r = stage_batch(keys, &d, &vb, false))) return r;
  if ((r = ensure_ws(ws2_, 16ull * n + 1024))) return r;
  Carver c(ws2_.p);
  uint64_t *d_hp = c.take<uint64_t>(n);
  uint64_t *d_addr = c.take<uint64_t>(n);
  int32_t *d_size = c.take<int32_t>(n);
  if (c.off > ws2_.cap && (r = ensure_ws(ws2_, c.off))) return r;
  GetArgs g{pools(), d.keys_hexs, d.keys_indexs, ntries_ > 1 ? d.trie_ids : nullptr, n, d.keys_bytes, ntries_, d_hp, d_addr, d_size};
  HIP_TRY(launch_gets(g, stream_));
  if (hps) HIP_TRY(hipMemcpyAsync(hps, d_hp, 8ull * n, hipMemcpyDeviceToHost, stream_));
  if (sizes) HIP_TRY(hipMemcpyAsync(sizes, d_size, 4ull * n, hipMemcpyDeviceToHost, stream_));
  return sync_counters();
}

int Forest::puts_with_read(const BatchDesc &b, const uint8_t *rw_flags, int mode, uint32_t *read_num,
                           uint64_t *read_hps, int32_t *read_sizes, void **targets, int *n_targets) {
  int r;
  if ((r = set_device())) return r;
  *read_num = 0;
  if (b.n == 0) return mode == 0 ? puts_2phase(b, targets, n_targets) : puts_latching(b, targets, n_targets);
  if (!rw_flags) {
    set_error("null rw_flags");
    return -3;
  }
  const uint32_t n = b.n;
  BatchDesc d;
  uint64_t vbase;
  if ((r = stage_batch(b, &d, &vbase, true))) return r;  // values land in the value arena (or stay adopted)
  // split layout: flags copy, isr, pos, nread, read keys/tids, write keys/values/hps/tids, scan temp
  const size_t need = n + 8ull * n + 16 + 12ull * n + 40ull * n + scan_temp_bytes(n) + 4096;
  if ((r = ensure_ws(rw_, need))) return r;
  Carver c(rw_.p);
  uint8_t *dflags = c.take<uint8_t>(n);
  uint32_t *isr = c.take<uint32_t>(n);
  uint32_t *pos = c.take<uint32_t>(n);
  uint32_t *nread = c.take<uint32_t>(4);
  int32_t *rk = c.take<int32_t>(2 * (size_t)n);
  int32_t *rtid = c.take<int32_t>(n);
  int32_t *wk = c.take<int32_t>(2 * (size_t)n);
  int64_t *wv = c.take<int64_t>(2 * (size_t)n);
  uint64_t *whp = c.take<uint64_t>(n);
  int32_t *wtid = c.take<int32_t>(n);
  void *temp = c.take<uint8_t>(scan_temp_bytes(n));
  if (c.off > rw_.cap) {
    set_error("workspace carve overflow");
    return -4;
  }
  HIP_TRY(hipMemcpyAsync(dflags, rw_flags, n, b.on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                         stream_));
  const int32_t *tid = ntries_ > 1 ? d.trie_ids : nullptr;
  HIP_TRY(rw_split(dflags, n, d.keys_indexs, d.values_indexs, d.values_hps, tid, isr, pos, nread, rk, rtid, wk, wv,
                   whp, wtid, &d_cnt_->err, temp, stream_));
  HIP_TRY(hipMemcpyAsync(h_scal_, nread, 4, hipMemcpyDeviceToHost, stream_));
  if ((r = sync_counters())) return r;
  const uint32_t nr = h_scal_[0];
  if (nr > n) {
    set_error("internal error: read count");
    return -4;
  }
  BatchDesc rb = d, wb = d;
  rb.keys_indexs = rk;
  rb.trie_ids = tid ? rtid : nullptr;
  rb.n = nr;
  rb.on_device = true;
  wb.keys_indexs = wk;
  wb.values_indexs = wv;
  wb.values_hps = d.values_hps ? whp : nullptr;
  wb.trie_ids = tid ? wtid : nullptr;
  wb.n = n - nr;
  wb.on_device = true;
  wb.adopt_values = true;  // already in the value arena, or the caller's adopted buffer
  wb.values_bytes = (const uint8_t *)(uintptr_t)vbase;
  if (nr > 0 && (r = gets(rb, read_hps, read_sizes))) return r;
  *read_num = nr;
  return mode == 0 ? puts_2phase(wb, targets, n_targets) : puts_latching(wb, targets, n_targets);
}

int Forest::get_values(const BatchDesc &keys, uint8_t *out, int64_t out_cap, int64_t *offsets, int32_t *sizes) {
  int r;
  if ((r = set_device())) return r;
  uint32_t n = keys.n;
  if (n == 0) return 0;
  BatchDesc d;
  uint64_t vb;
  if ((r = stage_batch(keys, &d, &vb, false))) return r;
  size_t need = 32ull * n + 4096;
  if ((r = ensure_ws(ws2_, need))) return r;
  Carver c(ws2_.p);
  uint64_t *d_hp = c.take<uint64_t>(n);
  uint64_t *d_addr = c.take<uint64_t>(n);
  int32_t *d_size = c.take<int32_t>(n);
  uint64_t *d_off = c.take<uint64_t>(n);
  GetArgs g{pools(), d.keys_hexs, d.keys_indexs, ntries_ > 1 ? d.trie_ids : nullptr, n, d.keys_bytes, ntries_, d_hp, d_addr, d_size};
  HIP_TRY(launch_gets(g, stream_));
  std::vector<int32_t> hs(n);
  HIP_TRY(hipMemcpyAsync(hs.data(), d_size, 4ull * n, hipMemcpyDeviceToHost, stream_));
  if ((r = sync_counters())) return r;
  std::vector<uint64_t> ho(n);
  uint64_t tot = 0;
  for (uint32_t i = 0; i < n; i++) {
    ho[i] = tot;
    if (hs[i] > 0) tot += (uint64_t)hs[i];
  }
  if ((int64_t)tot > out_cap) {
    set_error("get_values: output buffer too small (" + std::to_string(tot) + " bytes needed)");
    return -3;
  }
  DevBuf obuf;
  if (tot) {
    HIP_TRY(hipMalloc(&obuf.p, tot));
    HIP_TRY(hipMemcpyAsync(d_off, ho.data(), 8ull * n, hipMemcpyHostToDevice, stream_));
    HIP_TRY(launch_gather_values(d_addr, d_size, d_off, n, (uint8_t *)obuf.p, stream_));
    HIP_TRY(hipMemcpyAsync(out, obuf.p, tot, hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    HIP_TRY(hipFree(obuf.p));
  }
  for (uint32_t i = 0; i < n; i++) {
    if (offsets) offsets[i] = (int64_t)ho[i];
    if (sizes) sizes[i] = hs[i];
  }
  return 0;
}

int Forest::export_nodes(const BatchDesc &keys, uint8_t **hashes, uint8_t **encs, uint64_t **enc_indexs,
                         uint64_t *num) {
  int r;
  if ((r = set_device())) return r;
  *hashes = nullptr;
  *encs = nullptr;
  *enc_indexs = nullptr;
  *num = 0;
  uint32_t n = keys.n;
  if (n == 0) return 0;
  BatchDesc d;
  uint64_t vb;
  if ((r = stage_batch(keys, &d, &vb, false))) return r;
  return export_paths(this, d, hashes, encs, enc_indexs, num);
}

// LockNU.  Writes of the same key are combined first (the batch is sorted
// and only each key's last write is kept, which is the final state the
// sequential order gives), then the surviving writes insert concurrently with
// optimistic lock coupling.
int Forest::puts_latching(const BatchDesc &b, void **targets, int *n_targets) {
  int r;
  if ((r = set_device())) return r;
  if (b.n > 0x3fffffffu) {
    set_error("batch too large");
    return -3;
  }
  if (b.on_device && b.n && (b.keys_bytes == 0)) {
    set_error("device batch needs keys_bytes and values_total");
    return -3;
  }
  stats_.fallback_sorts = 0;
  stats_.regroups = 0;
  if (b.n > 0) {
    BatchDesc d;
    uint64_t vbase = 0;
    if ((r = stage_batch(b, &d, &vbase, true))) return r;
    const uint32_t n = b.n;
    all_empty_ = false;
    if ((r = ensure_pools(n, 2ull * n, n, d.key_slots))) return r;
    const bool has_tid = d.trie_ids != nullptr && ntries_ > 1;
    NuArgs a;
    void *temp = nullptr;
    uint32_t trie_bits = 0;
    if (has_tid) {
      while ((1ull << trie_bits) < ntries_) trie_bits++;
    }
    for (int run = 0;; run++) {
      if ((r = carve_nu(a, d, vbase, has_tid, &temp))) return r;
      a.seq_base = lock_seq_;
      a.skref = nullptr;  // the latched insert reads keys through kref only
      HIP_TRY(hipMemsetAsync(&d_cnt_->n_restarts, 0, 8, stream_));
      HIP_TRY(nu_prepare_keys(a, temp, stream_));
      int cur = 0;
      for (int pass = 0; pass >= 0;) {
        if ((r = sort_unique(a, trie_bits, temp, pass, &cur))) break;
        if ((r = sync_counters())) break;
        pass = next_pass(pass);
        if (pass >= 0) HIP_TRY(hipMemsetAsync(&d_cnt_->flag_fallback, 0, 4, stream_));
      }
      if (r == 0) break;
      if (run > 0 || !grow_key_arena_after_error()) return r;
    }
    const uint64_t ktot = *(const uint64_t *)(cnt_.scal + SC_KTOT);
    const uint32_t values_before = cnt_.n_value;
    HIP_TRY(lock_insert(a, key_used_ + ktot, temp, stream_));
    if ((r = sync_counters())) return r;
    // only new keys' nodes refer to the batch's packed keys: a batch that
    // only overwrites leaves the arena where it was
    if (cnt_.n_value != values_before) key_used_ = align_up(key_used_ + ktot, 8);
    lock_seq_ += n;
    stats_.restarts = cnt_.n_restarts;
    stats_.lock_fails = cnt_.n_lock_fail;
  }
  if (targets) *targets = dirty_.p;
  if (n_targets) *n_targets = (int)cnt_.n_dirty;
  return 0;
}

}  // namespace gmpt
