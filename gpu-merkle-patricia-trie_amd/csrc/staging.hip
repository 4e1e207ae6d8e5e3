// staging.hip -- pinned, chunked, multi-threaded host-to-device staging.
#include "staging.h"

#include <string.h>

#include <algorithm>

namespace gmpt {

WorkerPool::WorkerPool(unsigned workers) {
  for (unsigned i = 0; i < workers; i++) threads_.emplace_back(&WorkerPool::loop, this);
}

WorkerPool::~WorkerPool() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  for (auto &t : threads_) t.join();
}

void WorkerPool::loop() {
  uint64_t seen = 0;
  std::unique_lock<std::mutex> lk(mu_);
  for (;;) {
    cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
    if (stop_) return;
    seen = gen_;
    while (next_ < parts_) {
      const unsigned p = next_++;
      const std::function<void(unsigned)> *fn = fn_;
      lk.unlock();
      (*fn)(p);
      lk.lock();
      if (++finished_ == parts_) done_cv_.notify_all();
    }
  }
}

void WorkerPool::run(unsigned parts, const std::function<void(unsigned)> &fn) {
  if (parts == 0) return;
  std::unique_lock<std::mutex> lk(mu_);
  fn_ = &fn;
  parts_ = parts;
  next_ = 0;
  finished_ = 0;
  gen_++;
  cv_.notify_all();
  // the caller works too
  while (next_ < parts_) {
    const unsigned p = next_++;
    lk.unlock();
    fn(p);
    lk.lock();
    if (++finished_ == parts_) done_cv_.notify_all();
  }
  done_cv_.wait(lk, [&] { return finished_ == parts_; });
  parts_ = 0;
  fn_ = nullptr;
}

HostStager::~HostStager() {
  for (int k = 0; k < kBufs; k++) {
    if (ev_[k]) {
      (void)hipEventSynchronize(ev_[k]);
      (void)hipEventDestroy(ev_[k]);
    }
    if (pin_[k]) (void)hipHostFree(pin_[k]);
  }
  delete pool_;
}

hipError_t HostStager::init() {
  if (ready_) return hipSuccess;
  for (int k = 0; k < kBufs; k++) {
    hipError_t e = hipHostMalloc(&pin_[k], kChunk, hipHostMallocDefault);
    if (e != hipSuccess) return e;
    if ((e = hipEventCreateWithFlags(&ev_[k], hipEventDisableTiming)) != hipSuccess) return e;
  }
  // a few copy threads: more do not raise the memcpy rate into pinned pages
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  pool_ = new WorkerPool(std::min(5u, hw > 1 ? hw - 1 : 0u));
  ready_ = true;
  return hipSuccess;
}

hipError_t HostStager::copy(hipStream_t s, const StageSeg *segs, int nseg) {
  hipError_t e = init();
  if (e != hipSuccess) return e;
  int k = 0;
  for (int i = 0; i < nseg; i++) {
    const uint8_t *src = (const uint8_t *)segs[i].src;
    uint8_t *dst = (uint8_t *)segs[i].dst;
    for (size_t off = 0; off < segs[i].bytes; off += kChunk) {
      const size_t len = std::min(kChunk, segs[i].bytes - off);
      if (pending_[k]) {  // the buffer's previous DMA must be done before it is refilled
        if ((e = hipEventSynchronize(ev_[k])) != hipSuccess) return e;
        pending_[k] = false;
      }
      uint8_t *pin = (uint8_t *)pin_[k];
      const unsigned parts = pool_->size();
      const size_t per = ((len + parts - 1) / parts + 63) & ~(size_t)63;
      pool_->run(parts, [&](unsigned p) {
        const size_t a = std::min(len, (size_t)p * per), b = std::min(len, a + per);
        if (b > a) memcpy(pin + a, src + off + a, b - a);
      });
      if ((e = hipMemcpyAsync(dst + off, pin, len, hipMemcpyDefault, s)) != hipSuccess) return e;
      if ((e = hipEventRecord(ev_[k], s)) != hipSuccess) return e;
      pending_[k] = true;
      k = (k + 1) % kBufs;
    }
  }
  return hipSuccess;
}

void HostStager::extents(const int32_t *kidx, const int64_t *vidx, uint32_t n, uint64_t *kb, uint64_t *vb,
                         uint64_t *ks, int max_nibbles) {
  const unsigned parts = init() == hipSuccess ? pool_->size() * 2 : 1;
  std::vector<uint64_t> pk(parts, 0), pv(parts, 0), ps(parts, 0);
  auto body = [&](unsigned p) {
    const uint32_t a = (uint32_t)((uint64_t)n * p / parts), b = (uint32_t)((uint64_t)n * (p + 1) / parts);
    uint64_t k = 0, v = 0, sl = 0;
    for (uint32_t i = a; i < b; i++) {
      const int64_t st = kidx[2 * (size_t)i], en = kidx[2 * (size_t)i + 1];
      if (en >= 0 && (uint64_t)en + 1 > k) k = (uint64_t)en + 1;
      const int64_t L = (st >= 0 && en >= st && en - st < max_nibbles) ? en - st + 1 : 1;
      sl += (uint64_t)((L / 2 + 8) & ~7ll);
      if (vidx) {
        const int64_t ve = vidx[2 * (size_t)i + 1];
        if (ve >= 0 && (uint64_t)ve + 1 > v) v = (uint64_t)ve + 1;
      }
    }
    pk[p] = k;
    pv[p] = v;
    ps[p] = sl;
  };
  if (parts > 1)
    pool_->run(parts, body);
  else
    body(0);
  *kb = *vb = *ks = 0;
  for (unsigned p = 0; p < parts; p++) {
    *kb = std::max(*kb, pk[p]);
    *vb = std::max(*vb, pv[p]);
    *ks += ps[p];
  }
}

}  // namespace gmpt
