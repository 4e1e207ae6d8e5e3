// server.hip -- continuous-batching service (host runtime; see server.h).
#include "server.h"

#include <string.h>

#include <algorithm>
#include <chrono>
#include <unordered_set>

namespace gmpt {

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Admission estimate of the pool bytes one request can add: per key at most
// one branch, two short nodes and three value slots (the value node plus the
// slots PhaseNU reserves by position), its packed key and its value bytes.
uint64_t request_need(uint32_t n, uint64_t key_bytes, uint64_t value_bytes) {
  return (uint64_t)n * (sizeof(Branch) + 2 * sizeof(Short) + 3 * sizeof(Value) + 16) + key_bytes / 2 +
         value_bytes + 4096;
}

// The same checks the insert kernels apply, so a malformed request is refused
// at submit time instead of failing the batch it would share.
bool valid_batch(const BatchDesc &b, uint64_t kb, uint64_t vb, std::string *why) {
  for (uint32_t i = 0; i < b.n; i++) {
    int64_t s = b.keys_indexs[2 * (size_t)i], e = b.keys_indexs[2 * (size_t)i + 1];
    if (s < 0 || e < s || (uint64_t)e >= kb || e - s + 1 > kMaxKeyNibbles) {
      *why = "key " + std::to_string(i) + ": index pair out of range";
      return false;
    }
    for (int64_t j = s; j < e; j++)
      if (b.keys_hexs[j] > 15) {
        *why = "key " + std::to_string(i) + ": nibble above 15";
        return false;
      }
    if (b.keys_hexs[e] != kNibbleT) {
      *why = "key " + std::to_string(i) + ": missing terminator 16";
      return false;
    }
    int64_t vs = b.values_indexs[2 * (size_t)i], ve = b.values_indexs[2 * (size_t)i + 1];
    if (vs < 0 || ve < vs - 1 || (uint64_t)(ve + 1) > vb || ve - vs + 1 >= (1ll << 31)) {
      *why = "value " + std::to_string(i) + ": index pair out of range";
      return false;
    }
  }
  return true;
}

}  // namespace

int Server::Pinned::reserve(size_t bytes) {
  if (cap >= bytes) return 0;
  if (p) (void)hipHostFree(p);
  p = nullptr;
  cap = 0;
  size_t c = std::max<size_t>(bytes + bytes / 2, 1 << 16);
  if (hipHostMalloc(&p, c) != hipSuccess) {
    p = nullptr;
    set_error("pinned staging allocation failed");
    return -2;
  }
  cap = c;
  return 0;
}

Server::Pinned::~Pinned() {
  if (p) (void)hipHostFree(p);
}

Server::Server(int device, uint32_t ntries, uint32_t lanes, uint64_t mem_budget, uint32_t max_batch_keys, int mode)
    : device_(device), ntries_(ntries), nlanes_(lanes), max_batch_keys_(max_batch_keys), mode_(mode) {
  if (ntries == 0 || lanes == 0 || lanes > 16 || (mode != 0 && mode != 1)) {
    set_error("server: need ntries >= 1, 1 <= lanes <= 16, mode 0 or 1");
    return;
  }
  if (nlanes_ > ntries_) nlanes_ = ntries_;
  if (max_batch_keys_ == 0) max_batch_keys_ = 1u << 20;
  if (hipSetDevice(device) != hipSuccess) {
    set_error("server: hipSetDevice failed");
    return;
  }
  if (mem_budget == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      set_error("server: hipMemGetInfo failed");
      return;
    }
    mem_budget = (uint64_t)(fr * 0.8);
  }
  for (uint32_t l = 0; l < nlanes_; l++) {
    auto L = std::make_unique<Lane>();
    L->local_tries = (ntries_ - l + nlanes_ - 1) / nlanes_;
    L->f = std::make_unique<Forest>(device, L->local_tries);
    if (!L->f->ok()) return;
    L->budget = mem_budget / nlanes_;
    L->roots.resize(32 * (size_t)L->local_tries);
    L->root_sizes.resize(L->local_tries);
    lanes_.push_back(std::move(L));
  }
  for (uint32_t l = 0; l < nlanes_; l++) lanes_[l]->th = std::thread(&Server::worker, this, l);
  ok_ = true;
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  for (auto &L : lanes_)
    if (L->th.joinable()) L->th.join();
}

int Server::submit(uint32_t trie, const BatchDesc &b, uint64_t *id) {
  if (trie >= ntries_) {
    set_error("server: trie id out of range");
    return -3;
  }
  if (b.on_device) {
    set_error("server: requests are host batches");
    return -3;
  }
  if (b.n > 0 && (!b.keys_hexs || !b.keys_indexs || !b.values_bytes || !b.values_indexs)) {
    set_error("server: null batch array");
    return -3;
  }
  auto r = std::make_unique<Req>();
  r->trie = trie;
  r->n = b.n;
  uint64_t kb = 0, vb = 0;
  for (uint32_t i = 0; i < b.n; i++) {
    int64_t e = b.keys_indexs[2 * (size_t)i + 1], ve = b.values_indexs[2 * (size_t)i + 1];
    if (e >= 0) kb = std::max<uint64_t>(kb, (uint64_t)e + 1);
    if (ve >= 0) vb = std::max<uint64_t>(vb, (uint64_t)ve + 1);
  }
  std::string why;
  if (!valid_batch(b, kb, vb, &why)) {
    set_error("server: " + why);
    return -3;
  }
  r->keys.assign(b.keys_hexs, b.keys_hexs + kb);
  r->kidx.assign(b.keys_indexs, b.keys_indexs + 2 * (size_t)b.n);
  r->vals.assign(b.values_bytes, b.values_bytes + vb);
  r->vidx.assign(b.values_indexs, b.values_indexs + 2 * (size_t)b.n);
  r->need = request_need(b.n, kb, vb);
  uint32_t lane = trie % nlanes_;
  {
    std::lock_guard<std::mutex> g(mu_);
    if (r->need > lanes_[lane]->budget) {
      set_error("server: request exceeds the lane's memory budget");
      return -2;
    }
    r->t_submit = now_s();
    *id = reqs_.size();
    reqs_.push_back(std::move(r));
    lanes_[lane]->q.push_back(*id);
    st_.submitted++;
  }
  cv_work_.notify_all();
  return 0;
}

void Server::worker(uint32_t lane) {
  Lane &L = *lanes_[lane];
  if (hipSetDevice(device_) != hipSuccess) return;
  std::vector<Req *> batch;
  for (;;) {
    batch.clear();
    {
      std::unique_lock<std::mutex> g(mu_);
      cv_work_.wait(g, [&] { return stop_ || !L.q.empty(); });
      if (L.q.empty() && stop_) return;
      // Admit in arrival order: one request per trie, within the key cap and
      // the memory budget.  Once a trie is passed over, its later requests
      // wait too, so every trie completes its requests in order.
      std::unordered_set<uint32_t> seen;
      uint64_t used = L.f->pool_bytes_used(), admitted = 0, keys = 0;
      std::deque<uint64_t> rest;
      for (uint64_t id : L.q) {
        Req *r = reqs_[id].get();
        bool take = !seen.count(r->trie);
        seen.insert(r->trie);
        if (take && !batch.empty() && keys + r->n > max_batch_keys_) take = false;
        if (take && used + admitted + r->need > L.budget) {
          take = false;
          st_.deferred_mem++;
          if (batch.empty()) {
            // nothing runs on this lane that could free memory: refuse
            r->state = -1;
            r->err = "node pool budget exhausted";
            r->t_done = now_s();
            st_.failed++;
            cv_done_.notify_all();
            continue;
          }
        }
        if (take) {
          batch.push_back(r);
          admitted += r->need;
          keys += r->n;
          r->state = 1;
        } else {
          rest.push_back(id);
        }
      }
      L.q.swap(rest);
      if (batch.empty()) continue;
    }
    double t0 = now_s();
    int rc = run_batch(L, batch);
    double t1 = now_s();
    std::string err = rc ? last_error() : std::string();
    {
      std::lock_guard<std::mutex> g(mu_);
      for (Req *r : batch) {
        uint32_t lt = r->trie / nlanes_;
        if (rc == 0) {
          memcpy(r->root, &L.roots[32 * (size_t)lt], 32);
          r->root_size = L.root_sizes[lt];
          r->state = 2;
          st_.completed++;
        } else {
          r->state = -1;
          r->err = err;
          st_.failed++;
        }
        r->t_done = t1;
        // the request's bytes live in the trie now
        std::vector<uint8_t>().swap(r->keys);
        std::vector<int32_t>().swap(r->kidx);
        std::vector<uint8_t>().swap(r->vals);
        std::vector<int64_t>().swap(r->vidx);
      }
      st_.batches++;
      st_.batched_requests += batch.size();
      for (Req *r : batch) st_.batched_keys += r->n;
      st_.busy_s += t1 - t0;
    }
    cv_done_.notify_all();
  }
}

// Concatenates the admitted requests into one multi-trie host batch (pinned),
// inserts it, hashes, and reads the lane's roots.
int Server::run_batch(Lane &L, const std::vector<Req *> &batch) {
  size_t n = 0, kb = 0, vb = 0;
  for (Req *r : batch) {
    n += r->n;
    kb += r->keys.size();
    vb += r->vals.size();
  }
  int rc;
  if ((rc = L.keys.reserve(kb + 8)) || (rc = L.kidx.reserve(8 * n + 8)) || (rc = L.vals.reserve(vb + 8)) ||
      (rc = L.vidx.reserve(16 * n + 16)) || (rc = L.tid.reserve(4 * n + 4)))
    return rc;
  uint8_t *keys = (uint8_t *)L.keys.p, *vals = (uint8_t *)L.vals.p;
  int32_t *kidx = (int32_t *)L.kidx.p, *tid = (int32_t *)L.tid.p;
  int64_t *vidx = (int64_t *)L.vidx.p;
  size_t ko = 0, vo = 0, i = 0;
  for (Req *r : batch) {
    memcpy(keys + ko, r->keys.data(), r->keys.size());
    memcpy(vals + vo, r->vals.data(), r->vals.size());
    int32_t lt = (int32_t)(r->trie / nlanes_);
    for (uint32_t j = 0; j < r->n; j++, i++) {
      kidx[2 * i] = r->kidx[2 * j] + (int32_t)ko;
      kidx[2 * i + 1] = r->kidx[2 * j + 1] + (int32_t)ko;
      vidx[2 * i] = r->vidx[2 * j] + (int64_t)vo;
      vidx[2 * i + 1] = r->vidx[2 * j + 1] + (int64_t)vo;
      tid[i] = lt;
    }
    ko += r->keys.size();
    vo += r->vals.size();
  }
  if (ko > 0x7fffffffu) {
    set_error("server: batch key bytes exceed the int32 index range");
    return -3;
  }
  BatchDesc d;
  d.keys_hexs = keys;
  d.keys_indexs = kidx;
  d.values_bytes = vals;
  d.values_indexs = vidx;
  d.trie_ids = L.local_tries > 1 ? tid : nullptr;
  d.n = (uint32_t)n;
  void *targets = nullptr;
  int nt = 0;
  rc = mode_ == 0 ? L.f->puts_2phase(d, &targets, &nt) : L.f->puts_latching(d, &targets, &nt);
  if (rc) return rc;
  if ((rc = L.f->hash(targets, nt))) return rc;
  return L.f->root_hashes(0, L.local_tries, L.roots.data(), L.root_sizes.data());
}

int Server::poll(uint64_t id, int *state, uint8_t *root, int *root_size, double *latency_ms) {
  std::lock_guard<std::mutex> g(mu_);
  if (id >= reqs_.size()) {
    set_error("server: unknown request id");
    return -3;
  }
  Req *r = reqs_[id].get();
  *state = r->state;
  if (r->state == 2) {
    if (root) memcpy(root, r->root, 32);
    if (root_size) *root_size = r->root_size;
    if (latency_ms) *latency_ms = (r->t_done - r->t_submit) * 1e3;
  } else if (r->state == -1) {
    set_error("server: request failed: " + r->err);
  }
  return 0;
}

int Server::wait(uint64_t id, int *state, uint8_t *root, int *root_size, double *latency_ms) {
  {
    std::unique_lock<std::mutex> g(mu_);
    if (id >= reqs_.size()) {
      set_error("server: unknown request id");
      return -3;
    }
    Req *r = reqs_[id].get();
    cv_done_.wait(g, [&] { return r->state == 2 || r->state == -1; });
  }
  return poll(id, state, root, root_size, latency_ms);
}

int Server::drain() {
  std::unique_lock<std::mutex> g(mu_);
  cv_done_.wait(g, [&] { return st_.completed + st_.failed == st_.submitted; });
  return 0;
}

SrvStats Server::stats() {
  std::lock_guard<std::mutex> g(mu_);
  return st_;
}

std::vector<double> Server::latencies() {
  std::lock_guard<std::mutex> g(mu_);
  std::vector<double> out;
  for (auto &r : reqs_)
    if (r->state == 2) out.push_back((r->t_done - r->t_submit) * 1e3);
  return out;
}

}  // namespace gmpt
