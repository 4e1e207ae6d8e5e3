// capi.hip -- the library's C-ABI: gmpt_amd.h and the reference's libgmpt.h.
#include <stdlib.h>
#include <string.h>

#include <memory>
#include <new>

#include "../../include/gmpt_amd.h"
#include "../../include/libgmpt.h"
#include "engine.h"
#include "phasehc.h"
#include "prims.h"
#include "server.h"

using gmpt::BatchDesc;
using gmpt::Forest;

struct gmpt_forest {
  Forest *f = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
};

struct gmpt_server {
  gmpt::Server *s = nullptr;
};

namespace {

int fail(const char *msg) {
  gmpt::set_error(msg);
  return -3;
}

int to_desc(const gmpt_batch *b, BatchDesc *d) {
  if (!b) return fail("null batch");
  if (b->n < 0) return fail("negative batch size");
  if (b->n > 0 && (!b->keys_hexs || !b->keys_indexs)) return fail("batch without keys");
  d->keys_hexs = b->keys_hexs;
  d->keys_indexs = (const int32_t *)b->keys_indexs;
  d->values_bytes = b->values_bytes;
  d->values_indexs = b->values_indexs;
  d->values_hps = (const uint64_t *)b->values_hps;
  d->trie_ids = b->trie_ids;
  d->n = (uint32_t)b->n;
  d->keys_bytes = (uint64_t)(b->keys_bytes > 0 ? b->keys_bytes : 0);
  d->values_total = (uint64_t)(b->values_bytes_total > 0 ? b->values_bytes_total : 0);
  d->on_device = b->on_device != 0;
  d->adopt_values = b->adopt_values != 0;
  return 0;
}

int check_put_batch(const gmpt_batch *b) {
  if (b->n > 0 && (!b->values_bytes || !b->values_indexs)) return fail("batch without values");
  if (b->n > 0 && b->on_device && (b->keys_bytes <= 0 || b->values_bytes_total < 0))
    return fail("device batch needs keys_bytes and values_bytes_total");
  return 0;
}

}  // namespace

extern "C" {

int gmpt_forest_create(int device, uint32_t num_tries, gmpt_forest **out) {
  if (!out) return fail("null output");
  *out = nullptr;
  if (num_tries == 0) return fail("a forest needs at least one trie");
  std::unique_ptr<gmpt_forest> h(new (std::nothrow) gmpt_forest);
  if (!h) return fail("out of host memory");
  h->f = new (std::nothrow) Forest(device, num_tries);
  if (!h->f || !h->f->ok()) {
    delete h->f;
    return fail(gmpt::last_error()[0] ? gmpt::last_error() : "device initialisation failed");
  }
  if (hipEventCreate(&h->t0) != hipSuccess || hipEventCreate(&h->t1) != hipSuccess) {
    delete h->f;
    return fail("hipEventCreate failed");
  }
  *out = h.release();
  return 0;
}

void gmpt_forest_destroy(gmpt_forest *f) {
  if (!f) return;
  delete f->f;
  if (f->t0) (void)hipEventDestroy(f->t0);
  if (f->t1) (void)hipEventDestroy(f->t1);
  delete f;
}

const char *gmpt_last_error(void) { return gmpt::last_error(); }

int gmpt_puts_2phase(gmpt_forest *f, const gmpt_batch *b, void **targets, int *n_targets) {
  if (!f) return fail("null forest");
  BatchDesc d;
  int r;
  if ((r = to_desc(b, &d)) || (r = check_put_batch(b))) return r;
  return f->f->puts_2phase(d, targets, n_targets);
}

int gmpt_rebuild_graph(gmpt_forest *f, const gmpt_batch *b, int *replayed) {
  if (!f || !replayed) return fail("null argument");
  BatchDesc d;
  int r;
  if ((r = to_desc(b, &d)) || (r = check_put_batch(b))) return r;
  bool rep = false;
  r = f->f->rebuild_graph(d, &rep);
  *replayed = rep ? 1 : 0;
  return r;
}

int gmpt_puts_latching(gmpt_forest *f, const gmpt_batch *b, void **targets, int *n_targets) {
  if (!f) return fail("null forest");
  BatchDesc d;
  int r;
  if ((r = to_desc(b, &d)) || (r = check_put_batch(b))) return r;
  return f->f->puts_latching(d, targets, n_targets);
}

int gmpt_puts_with_read(gmpt_forest *f, const gmpt_batch *b, const uint8_t *rw_flags, int mode, int *read_num,
                        const uint8_t **read_values_hps, int *read_values_sizes, void **targets, int *n_targets) {
  if (!f || !read_num) return fail("null argument");
  if (mode != 0 && mode != 1) return fail("mode must be 0 (PhaseNU) or 1 (LockNU)");
  BatchDesc d;
  int r;
  if ((r = to_desc(b, &d)) || (r = check_put_batch(b))) return r;
  std::unique_ptr<uint64_t[]> hp(new (std::nothrow) uint64_t[d.n + 1]);
  std::unique_ptr<int32_t[]> sz(new (std::nothrow) int32_t[d.n + 1]);
  if (!hp || !sz) return fail("out of host memory");
  uint32_t nr = 0;
  r = f->f->puts_with_read(d, rw_flags, mode, &nr, hp.get(), sz.get(), targets, n_targets);
  *read_num = (int)nr;
  if (r) return r;
  for (uint32_t i = 0; i < nr; i++) {
    bool found = sz[i] >= 0;
    if (read_values_hps) read_values_hps[i] = found ? (const uint8_t *)(uintptr_t)hp[i] : nullptr;
    if (read_values_sizes) read_values_sizes[i] = found ? sz[i] : 0;
  }
  return 0;
}

int gmpt_hash_onepass(gmpt_forest *f, void *targets, int n_targets) {
  if (!f) return fail("null forest");
  return f->f->hash(targets, n_targets);
}

int gmpt_get_root_hash(gmpt_forest *f, uint32_t trie, uint8_t out[32], int *size) {
  if (!f || !size) return fail("null argument");
  return f->f->root_hash(trie, out, size);
}

int gmpt_get_root_hashes(gmpt_forest *f, uint32_t first, uint32_t count, uint8_t *out, int *sizes) {
  if (!f || (count && (!out || !sizes))) return fail("null argument");
  return f->f->root_hashes(first, count, out, sizes);
}

int gmpt_gets(gmpt_forest *f, const gmpt_batch *keys, const uint8_t **values_hps, int *values_sizes) {
  if (!f) return fail("null forest");
  BatchDesc d;
  int r;
  if ((r = to_desc(keys, &d))) return r;
  if (d.n == 0) return 0;
  std::unique_ptr<uint64_t[]> hp(new uint64_t[d.n]);
  std::unique_ptr<int32_t[]> sz(new int32_t[d.n]);
  if ((r = f->f->gets(d, hp.get(), sz.get()))) return r;
  for (uint32_t i = 0; i < d.n; i++) {
    bool found = sz[i] >= 0;
    if (values_hps) values_hps[i] = found ? (const uint8_t *)(uintptr_t)hp[i] : nullptr;
    if (values_sizes) values_sizes[i] = found ? sz[i] : 0;
  }
  return 0;
}

int gmpt_get_values(gmpt_forest *f, const gmpt_batch *keys, uint8_t *out, int64_t out_cap, int64_t *offsets,
                    int *sizes) {
  if (!f) return fail("null forest");
  BatchDesc d;
  int r;
  if ((r = to_desc(keys, &d))) return r;
  return f->f->get_values(d, out, out_cap, offsets, (int32_t *)sizes);
}

int gmpt_export_nodes(gmpt_forest *f, const gmpt_batch *keys, const uint8_t **hashes, const uint8_t **encs,
                      const uint64_t **enc_indexs, uint64_t *num) {
  if (!f || !hashes || !encs || !enc_indexs || !num) return fail("null argument");
  BatchDesc d;
  int r;
  if ((r = to_desc(keys, &d))) return r;
  uint8_t *h = nullptr, *e = nullptr;
  uint64_t *ix = nullptr;
  r = f->f->export_nodes(d, &h, &e, &ix, num);
  *hashes = h;
  *encs = e;
  *enc_indexs = ix;
  return r;
}

void gmpt_free_host(const void *p) { delete[] (const uint8_t *)p; }

int gmpt_reset_trie(gmpt_forest *f, uint32_t trie) {
  if (!f) return fail("null forest");
  return f->f->reset_trie(trie);
}

int gmpt_forest_clear(gmpt_forest *f) {
  if (!f) return fail("null forest");
  return f->f->clear();
}

int gmpt_set_option(gmpt_forest *f, const char *name, int64_t value) {
  if (!f || !name) return fail("null argument");
  return f->f->set_option(name, value);
}

int gmpt_get_option(gmpt_forest *f, const char *name, int64_t *value) {
  if (!f || !name || !value) return fail("null argument");
  return f->f->get_option(name, value);
}

int gmpt_autotune(gmpt_forest *f, const gmpt_batch *sample, int mode, int64_t *chosen, double *best_ms) {
  if (!f) return fail("null forest");
  if (mode != 0 && mode != 1) return fail("mode must be 0 (PhaseNU) or 1 (LockNU)");
  BatchDesc d;
  int r;
  if ((r = to_desc(sample, &d)) || (r = check_put_batch(sample))) return r;
  return f->f->autotune(d, mode, chosen, best_ms);
}

int gmpt_synchronize(gmpt_forest *f) {
  if (!f) return fail("null forest");
  return f->f->synchronize();
}

int gmpt_get_stats(gmpt_forest *f, gmpt_stats *out) {
  if (!f || !out) return fail("null argument");
  const gmpt::OpStats &s = f->f->stats();
  memset(out, 0, sizeof(*out));
  out->last_hash_ms = s.last_hash_ms;
  out->hash_launches = s.hash_launches;
  out->hashed_nodes = s.hashed_nodes;
  out->restarts = s.restarts;
  out->lock_fails = s.lock_fails;
  out->fallback_sorts = s.fallback_sorts;
  out->regroups = s.regroups;
  out->perms = s.perms;
  out->enc_bytes = s.enc_bytes;
  out->leaf_perm_ms = s.leaf_perm_ms;
  out->leaf_perms = s.leaf_perms;
  out->leaf_bytes = s.leaf_bytes;
  out->leaf_nodes = s.leaf_nodes;
  out->ext_direct_nodes = s.ext_direct_nodes;
  out->pool_bytes_used = f->f->pool_bytes_used();
  out->pool_bytes_reserved = f->f->pool_bytes_reserved();
  return 0;
}

#define CAPI_TRY(expr)                                                 \
  do {                                                                 \
    hipError_t e__ = (expr);                                           \
    if (e__ != hipSuccess) {                                           \
      gmpt::set_error(std::string(#expr) + ": " + hipGetErrorString(e__)); \
      return -1;                                                       \
    }                                                                  \
  } while (0)

int gmpt_device_count(int *count) {
  CAPI_TRY(hipGetDeviceCount(count));
  return 0;
}
int gmpt_device_malloc(void **p, size_t bytes) {
  CAPI_TRY(hipMalloc(p, bytes ? bytes : 1));
  return 0;
}
int gmpt_device_free(void *p) {
  CAPI_TRY(hipFree(p));
  return 0;
}
int gmpt_memcpy_htod(void *dst, const void *src, size_t bytes) {
  CAPI_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return 0;
}
int gmpt_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  CAPI_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return 0;
}

int gmpt_device_copy(void *dst, const void *src, size_t bytes) {
  if (bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16) return fail("copy needs 16-byte aligned sizes");
  CAPI_TRY(gmpt::device_copy16(dst, src, bytes, nullptr));
  CAPI_TRY(hipDeviceSynchronize());
  return 0;
}

int gmpt_timer_start(gmpt_forest *f) {
  if (!f) return fail("null forest");
  CAPI_TRY(hipEventRecord(f->t0, f->f->stream()));
  return 0;
}

int gmpt_timer_stop(gmpt_forest *f, float *ms) {
  if (!f || !ms) return fail("null argument");
  CAPI_TRY(hipEventRecord(f->t1, f->f->stream()));
  CAPI_TRY(hipEventSynchronize(f->t1));
  CAPI_TRY(hipEventElapsedTime(ms, f->t0, f->t1));
  return 0;
}

int gmpt_keccak256_batch_mode(const uint8_t *data, const uint64_t *offsets, const uint32_t *lengths, uint32_t n,
                              uint8_t *out, int mode) {
  if (n == 0) return 0;
  if (!data || !offsets || !lengths || !out) return fail("null argument");
  if (mode != 0 && mode != 1) return fail("keccak mode must be 0 (lane) or 1 (spread)");
  uint64_t total = 0;
  for (uint32_t i = 0; i < n; i++) {
    uint64_t e = offsets[i] + lengths[i];
    if (e > total) total = e;
  }
  uint64_t sbytes = gmpt::keccak_scratch_bytes(lengths, n);
  uint8_t *d_data = nullptr, *d_out = nullptr;
  uint64_t *d_off = nullptr, *d_slot = nullptr, *d_scratch = nullptr;
  uint32_t *d_len = nullptr;
  void *d_temp = nullptr;
  int rc = 0;
  hipError_t e;
  if ((e = hipMalloc(&d_data, total + 16)) != hipSuccess || (e = hipMalloc(&d_off, 8ull * n)) != hipSuccess ||
      (e = hipMalloc(&d_len, 4ull * n)) != hipSuccess || (e = hipMalloc(&d_out, 32ull * n)) != hipSuccess ||
      (e = hipMalloc(&d_slot, 8ull * n)) != hipSuccess || (e = hipMalloc(&d_scratch, sbytes + 64)) != hipSuccess ||
      (e = hipMalloc(&d_temp, gmpt::scan_temp_bytes(n))) != hipSuccess ||
      (e = hipMemcpy(d_data, data, total, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(d_off, offsets, 8ull * n, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(d_len, lengths, 4ull * n, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = gmpt::keccak_batch(d_data, d_off, d_len, n, d_out, d_slot, d_scratch, d_temp, mode, nullptr)) !=
          hipSuccess ||
      (e = hipMemcpy(out, d_out, 32ull * n, hipMemcpyDeviceToHost)) != hipSuccess) {
    gmpt::set_error(std::string("keccak batch: ") + hipGetErrorString(e));
    rc = -1;
  }
  void *bufs[] = {d_data, d_off, d_len, d_out, d_slot, d_scratch, d_temp};
  for (void *p : bufs)
    if (p) (void)hipFree(p);
  return rc;
}

int gmpt_keccak256_batch(const uint8_t *data, const uint64_t *offsets, const uint32_t *lengths, uint32_t n,
                         uint8_t *out) {
  return gmpt_keccak256_batch_mode(data, offsets, lengths, n, out, 0);
}

// ------------------------------------------------------------------ server --

int gmpt_server_create(int device, uint32_t num_tries, uint32_t lanes, uint64_t mem_budget, uint32_t max_batch_keys,
                       int mode, gmpt_server **out) {
  if (!out) return fail("null output");
  *out = nullptr;
  std::unique_ptr<gmpt::Server> s(new (std::nothrow) gmpt::Server(device, num_tries, lanes, mem_budget,
                                                                   max_batch_keys, mode));
  if (!s) return fail("out of host memory");
  if (!s->ok()) return -1;
  gmpt_server *h = new (std::nothrow) gmpt_server;
  if (!h) return fail("out of host memory");
  h->s = s.release();
  *out = h;
  return 0;
}

void gmpt_server_destroy(gmpt_server *s) {
  if (!s) return;
  delete s->s;
  delete s;
}

int gmpt_server_submit(gmpt_server *s, uint32_t trie, const gmpt_batch *b, uint64_t *req_id) {
  if (!s || !req_id) return fail("null argument");
  BatchDesc d;
  int r;
  if ((r = to_desc(b, &d)) || (r = check_put_batch(b))) return r;
  return s->s->submit(trie, d, req_id);
}

int gmpt_server_poll(gmpt_server *s, uint64_t req_id, int *state, uint8_t root[32], int *root_size,
                     double *latency_ms) {
  if (!s || !state) return fail("null argument");
  return s->s->poll(req_id, state, root, root_size, latency_ms);
}

int gmpt_server_wait(gmpt_server *s, uint64_t req_id, int *state, uint8_t root[32], int *root_size,
                     double *latency_ms) {
  if (!s || !state) return fail("null argument");
  return s->s->wait(req_id, state, root, root_size, latency_ms);
}

int gmpt_server_drain(gmpt_server *s) {
  if (!s) return fail("null server");
  return s->s->drain();
}

int gmpt_server_get_stats(gmpt_server *s, gmpt_server_stats *out) {
  if (!s || !out) return fail("null argument");
  gmpt::SrvStats st = s->s->stats();
  out->submitted = st.submitted;
  out->completed = st.completed;
  out->failed = st.failed;
  out->batches = st.batches;
  out->batched_requests = st.batched_requests;
  out->batched_keys = st.batched_keys;
  out->deferred_mem = st.deferred_mem;
  out->busy_s = st.busy_s;
  return 0;
}

int gmpt_server_latencies(gmpt_server *s, double *out, uint64_t cap, uint64_t *count) {
  if (!s || !count) return fail("null argument");
  std::vector<double> v = s->s->latencies();
  *count = v.size();
  if (out) memcpy(out, v.data(), sizeof(double) * std::min<uint64_t>(cap, v.size()));
  return 0;
}

// ------------------------------------------------------------- libgmpt.h --
// One forest holds the three tries geth builds per block (state,
// transaction, receipt), trie id = TrieType.

static gmpt_forest *g_tries = nullptr;

static int env_device() {
  const char *s = getenv("GMPT_DEVICE");
  return s ? atoi(s) : 0;
}

int preprocess() {
  if (g_tries) {
    gmpt_forest_destroy(g_tries);
    g_tries = nullptr;
  }
  return gmpt_forest_create(env_device(), 3, &g_tries) == 0 ? 1 : 0;
}

static const uint8_t *build_common(gmpt_forest *f, uint32_t trie, const uint8_t *keys_hexs, int *keys_hexs_indexs,
                                   const uint8_t *values_bytes, int64_t *values_bytes_indexs,
                                   const uint8_t **values_hps, int insert_num, bool latching) {
  if (!f) {
    gmpt::set_error("trie not initialised (call preprocess / init_mpt first)");
    return nullptr;
  }
  std::unique_ptr<int32_t[]> tid;
  gmpt_batch b;
  memset(&b, 0, sizeof(b));
  b.keys_hexs = keys_hexs;
  b.keys_indexs = keys_hexs_indexs;
  b.values_bytes = values_bytes;
  b.values_indexs = values_bytes_indexs;
  b.values_hps = values_hps;
  b.n = insert_num;
  if (f->f->ntries() > 1 && insert_num > 0) {
    tid.reset(new int32_t[insert_num]);
    for (int i = 0; i < insert_num; i++) tid[i] = (int32_t)trie;
    b.trie_ids = tid.get();
  }
  void *targets = nullptr;
  int nt = 0;
  int r = latching ? gmpt_puts_latching(f, &b, &targets, &nt) : gmpt_puts_2phase(f, &b, &targets, &nt);
  if (r) return nullptr;
  if (gmpt_hash_onepass(f, targets, nt)) return nullptr;
  uint8_t *h = new uint8_t[32]();
  int hs = 0;
  if (gmpt_get_root_hash(f, trie, h, &hs) || hs == 0) {
    delete[] h;
    return nullptr;
  }
  return h;
}

const uint8_t *build_mpt_2phase(enum TrieType trie_type, const uint8_t *keys_hexs, int *keys_hexs_indexs,
                                const uint8_t *values_bytes, int64_t *values_bytes_indexs,
                                const uint8_t **values_hps, int insert_num) {
  return build_common(g_tries, (uint32_t)trie_type, keys_hexs, keys_hexs_indexs, values_bytes, values_bytes_indexs,
                      values_hps, insert_num, false);
}

const uint8_t *build_mpt_olc(enum TrieType trie_type, const uint8_t *keys_hexs, int *keys_hexs_indexs,
                             const uint8_t *values_bytes, int64_t *values_bytes_indexs, const uint8_t **values_hps,
                             int insert_num) {
  return build_common(g_tries, (uint32_t)trie_type, keys_hexs, keys_hexs_indexs, values_bytes, values_bytes_indexs,
                      values_hps, insert_num, true);
}

struct nodeset get_all_nodes(enum TrieType trie_type, const uint8_t *keys_hexs, int *keys_hexs_indexs, int num) {
  struct nodeset set;
  memset(&set, 0, sizeof(set));
  if (!g_tries) {
    gmpt::set_error("trie not initialised (call preprocess first)");
    return set;
  }
  std::unique_ptr<int32_t[]> tid(new int32_t[num > 0 ? num : 1]);
  for (int i = 0; i < num; i++) tid[i] = (int32_t)trie_type;
  gmpt_batch b;
  memset(&b, 0, sizeof(b));
  b.keys_hexs = keys_hexs;
  b.keys_indexs = keys_hexs_indexs;
  b.trie_ids = tid.get();
  b.n = num;
  const uint8_t *h = nullptr, *e = nullptr;
  const uint64_t *ix = nullptr;
  uint64_t n = 0;
  if (gmpt_export_nodes(g_tries, &b, &h, &e, &ix, &n)) return set;
  set.hashs = h;
  set.encs = e;
  set.encs_indexs = (const unsigned long long *)ix;
  set.num = n;
  return set;
}

void *init_mpt() {
  gmpt_forest *f = nullptr;
  if (gmpt_forest_create(env_device(), 1, &f)) return nullptr;
  return f;
}

const uint8_t *insert_mpt_2phase(void *mpt, const uint8_t *keys_hexs, int *keys_hexs_indexs,
                                 const uint8_t *values_bytes, int64_t *values_bytes_indexs,
                                 const uint8_t **values_hps, int insert_num) {
  return build_common((gmpt_forest *)mpt, 0, keys_hexs, keys_hexs_indexs, values_bytes, values_bytes_indexs,
                      values_hps, insert_num, false);
}

const uint8_t *insert_mpt_olc(void *mpt, const uint8_t *keys_hexs, int *keys_hexs_indexs,
                              const uint8_t *values_bytes, int64_t *values_bytes_indexs,
                              const uint8_t **values_hps, int insert_num) {
  return build_common((gmpt_forest *)mpt, 0, keys_hexs, keys_hexs_indexs, values_bytes, values_bytes_indexs,
                      values_hps, insert_num, true);
}

void free_mpt(void *mpt) { gmpt_forest_destroy((gmpt_forest *)mpt); }

}  // extern "C"
