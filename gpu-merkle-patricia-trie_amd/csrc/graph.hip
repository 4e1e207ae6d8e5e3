// graph.hip -- a bulk rebuild pass (clear + PhaseNU + PhaseHC) captured once
// as a HIP graph and replayed per batch.
//
// An eager pass synchronises with the host four times: after the sort (to
// see whether a fallback sort is needed), after the build (counters), after
// the hash classification (the host reads the level sizes to launch one
// range per level) and at the end.  A replay runs the same launches with
// the level sizes of the captured pass and one synchronisation at the end.
// It is valid only when the batch behaves like the captured one, which the
// device checks itself right after the classification (k_graph_check): no
// sort fallback, no error, the same dirty-list length, element count and
// hash bins.  Otherwise the hash kernels skip (the levels' ranges would be
// wrong) and the host redoes the pass eagerly and captures it again.
#include <string.h>

#include "engine.h"
#include "phasenu.h"

namespace gmpt {

#define GRAPH_TRY(expr)                                                            \
  do {                                                                             \
    hipError_t e__ = (expr);                                                       \
    if (e__ != hipSuccess) {                                                       \
      set_error(std::string(#expr) + ": " + hipGetErrorString(e__));              \
      return -1;                                                                   \
    }                                                                              \
  } while (0)

void Forest::drop_graph() {
  if (gexec_) (void)hipGraphExecDestroy(gexec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  gexec_ = nullptr;
  graph_ = nullptr;
  plan_.valid = false;
}

static bool same_batch(const BatchDesc &a, const BatchDesc &b) {
  return a.keys_hexs == b.keys_hexs && a.keys_indexs == b.keys_indexs && a.values_bytes == b.values_bytes &&
         a.values_indexs == b.values_indexs && a.values_hps == b.values_hps && a.trie_ids == b.trie_ids &&
         a.n == b.n && a.keys_bytes == b.keys_bytes && a.values_total == b.values_total;
}

// clear + PhaseNU + PhaseHC as the plain calls run them; records the pass's
// shape for a capture.
int Forest::rebuild_eager(const BatchDesc &b) {
  int r;
  void *targets = nullptr;
  int nt = 0;
  if ((r = clear())) return r;
  if ((r = puts_2phase(b, &targets, &nt))) return r;
  plan_.n_dirty = cnt_.n_dirty;
  plan_.ne = cnt_.scal[SC_NE];
  plan_.key_used = key_used_;
  if ((r = hash(targets, nt))) return r;
  if (plan_.n_dirty) memcpy(plan_.hist, ((const HcMeta *)h_meta_)->hist, sizeof(plan_.hist));
  return 0;
}

// Captures the pass the eager run just made (every buffer it needs is sized
// already, so no allocation or synchronisation happens while capturing).
int Forest::capture_rebuild(const BatchDesc &b) {
  int r;
  if ((r = ensure_ws(plan_hist_, sizeof(plan_.hist)))) return r;
  GRAPH_TRY(hipMemcpyAsync(plan_hist_.p, plan_.hist, sizeof(plan_.hist), hipMemcpyHostToDevice, stream_));
  GRAPH_TRY(hipStreamSynchronize(stream_));
  // the host side of clear(): the graph's first kernel clears the device side
  memset(&cnt_, 0, sizeof(cnt_));
  root0_ok_ = false;
  key_used_ = 0;
  vchunk_used_ = 0;
  rec_ = LeafRecords();
  BatchDesc d;
  uint64_t vbase = 0;
  if ((r = stage_batch(b, &d, &vbase, true))) return r;
  const bool has_tid = d.trie_ids != nullptr && ntries_ > 1;
  uint32_t trie_bits = 0;
  if (has_tid)
    while ((1ull << trie_bits) < ntries_) trie_bits++;
  GRAPH_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  hipGraph_t g = nullptr;
  auto fail = [&](int code) {
    (void)hipStreamEndCapture(stream_, &g);
    if (g) (void)hipGraphDestroy(g);
    return code;
  };
  hipError_t e = hipSuccess;
  Pools P = pools();
  if ((e = hc_clear(d_cnt_, (TrieHdr *)tries_.p, ntries_, stream_)) != hipSuccess) return fail(-1);
  NuArgs a;
  void *temp = nullptr;
  if ((r = carve_nu(a, d, vbase, has_tid, &temp))) return fail(r);
  a.e_ident = has_tid ? 0u : 1u;  // the pass builds into emptied tries
  uint64_t *sk = nullptr;
  uint32_t *sv = nullptr;
  uint64_t sk_mask = 0;
  int cur = 0;
  if ((e = nu_prepare_keys(a, temp, stream_)) != hipSuccess ||
      (e = nu_sort(a, 0, trie_bits, temp, stream_, &sk, &sv, false, &cur, 0, &sk_mask)) != hipSuccess)
    return fail(-1);
  a.u_win_ok = sk_mask == 0;
  if ((e = nu_adjacent(a, sk, sv, 0, stream_, sk_mask)) != hipSuccess ||
      (e = nu_unique(a, sk, sv, temp, stream_)) != hipSuccess ||
      (e = nu_descend_and_group(a, temp, stream_, true)) != hipSuccess ||
      (e = nu_emit(a, temp, stream_, true)) != hipSuccess || (e = nu_build(a, temp, stream_, true)) != hipSuccess)
    return fail(-1);
  // the hash, with the captured pass's sizes
  const uint32_t n = plan_.n_dirty;
  unsigned long long *parts = nullptr;
  if (n > 0) {
    uint64_t *lmeta = hc_fuse_leaves_ ? (uint64_t *)hc_lmeta_.p : nullptr;
    uint64_t *lmeta_tmp = hc_fuse_leaves_ ? (uint64_t *)hc_lmeta_tmp_.p : nullptr;
    LeafRecords rec;
    rec.pk = a.leaf_pk;
    rec.meta = a.leaf_meta;
    rec.lo = 0;  // the dirty list starts empty after the clear
    rec.n = plan_.ne;
    // then the build's branches and their extension slots, as many each
    rec.bdesc = a.bdesc;
    rec.blo = plan_.ne;
    rec.bn = (plan_.n_dirty - std::min(plan_.n_dirty, plan_.ne)) / 2;
    HcMeta *meta = (HcMeta *)hc_meta_.p;
    if ((e = hc_classify(P, (const ref_t *)dirty_.p, n, meta, (uint32_t *)hc_cursor_.p, (ref_t *)hc_sorted_.p,
                         (uint64_t *)hc_slot_.p, (uint32_t *)hc_enc_.p, ws2_.p, stream_, lmeta_tmp, lmeta, rec)) !=
            hipSuccess ||
        (e = hc_graph_check(P, meta, (const uint32_t *)plan_hist_.p, plan_.n_dirty, plan_.ne, a.scal, stream_)) !=
            hipSuccess)
      return fail(-1);
    if (lmeta && (e = hc_hash_leaves_direct(P, (const ref_t *)hc_sorted_.p, lmeta, meta, n, (uint32_t *)hc_enc_.p,
                                            stream_)) != hipSuccess)
      return fail(-1);
    uint32_t n_leaf = 0;
    stats_.hash_launches = stats_.hashed_nodes = stats_.ext_direct_nodes = 0;
    if ((r = hash_levels(P, plan_.hist, lmeta != nullptr, true, &n_leaf))) return fail(r);
    stats_.leaf_nodes = n_leaf;
    parts = (unsigned long long *)((uint8_t *)hc_meta_.p + sizeof(HcMeta));
    if ((e = hc_count((const uint32_t *)hc_enc_.p, meta, n, n_leaf, parts, stream_)) != hipSuccess) return fail(-1);
  }
  if ((e = hc_roots(P, ntries_, parts, stream_)) != hipSuccess ||
      (e = hipMemcpyAsync(h_cnt_, d_cnt_, sizeof(Counters), hipMemcpyDeviceToHost, stream_)) != hipSuccess)
    return fail(-1);
  GRAPH_TRY(hipStreamEndCapture(stream_, &g));
  hipGraphExec_t x = nullptr;
  if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(g);
    set_error("hipGraphInstantiate failed");
    return -1;
  }
  graph_ = g;
  gexec_ = x;
  plan_.batch = b;
  plan_.valid = true;
  return 0;
}

int Forest::rebuild_graph(const BatchDesc &b, bool *replayed) {
  int r;
  if ((r = set_device())) return r;
  *replayed = false;
  if (arena_) {
    set_error("a forest over a fixed device region does not capture graphs");
    return -3;
  }
  if (!b.on_device || !b.adopt_values || b.n == 0) {
    set_error("a graph rebuild takes a non-empty device batch whose values the forest adopts");
    return -3;
  }
  if (plan_.valid && same_batch(plan_.batch, b)) {
    if ((r = replay_graph())) return r;
    if ((*replayed = graph_ok())) return 0;
    // unlike the captured batch: redo it eagerly (from a clear) and capture that
  }
  drop_graph();
  if ((r = rebuild_eager(b))) return r;
  // capture the pass just run, then replay it once: its result replaces the
  // eager one (the capture reset the host mirrors, the replay refreshes them)
  if (capture_rebuild(b) == 0) {
    if ((r = replay_graph())) return r;
    if (graph_ok()) return 0;
  }
  // a capture or replay that fails leaves no graph; the eager pass stands
  drop_graph();
  return rebuild_eager(b);
}

int Forest::replay_graph() {
  GRAPH_TRY(hipGraphLaunch(gexec_, stream_));
  GRAPH_TRY(hipStreamSynchronize(stream_));
  return 0;
}

// After a replay: whether it was valid; if so the host mirrors take its counters.
bool Forest::graph_ok() {
  const Counters c = *h_cnt_;
  if (c.err != ERR_NONE || c.graph_bad != 0 || c.flag_fallback != 0) return false;
  cnt_ = c;
  root0_ok_ = c.root0_size <= 32;  // the replay's counter read follows its root pass
  key_used_ = plan_.key_used;
  all_empty_ = false;
  rec_ = LeafRecords();
  stats_.fallback_sorts = stats_.regroups = 0;
  stats_.perms = cnt_.n_perms;
  stats_.enc_bytes = cnt_.enc_bytes;
  stats_.leaf_perms = cnt_.leaf_perms;
  stats_.leaf_bytes = cnt_.leaf_bytes;
  stats_.last_hash_ms = stats_.leaf_perm_ms = 0;
  return true;
}

}  // namespace gmpt
