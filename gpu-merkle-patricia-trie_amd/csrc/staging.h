// staging.h -- host-to-device staging of host batches through pinned memory.
//
// A host batch arrives in pageable memory, which the DMA engines cannot read
// directly: a plain hipMemcpyAsync copies it through the runtime's own pinned
// bounce buffer with one CPU thread (about 22 GB/s for a 1M-key batch here).
// The stager instead splits every array into chunks, has a few worker
// threads copy each chunk into one of a ring of pinned buffers, and queues
// the DMA of that buffer on the forest's stream while the workers fill the
// next one, so CPU copies and PCIe transfers overlap.  The same workers scan
// the batch's index pairs for the buffer extents.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace gmpt {

// A fixed set of worker threads that run fn(0 .. parts - 1) together with the
// calling thread.
class WorkerPool {
 public:
  explicit WorkerPool(unsigned workers);
  ~WorkerPool();
  unsigned size() const { return (unsigned)threads_.size() + 1; }
  void run(unsigned parts, const std::function<void(unsigned)> &fn);

 private:
  void loop();
  std::vector<std::thread> threads_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(unsigned)> *fn_ = nullptr;
  unsigned parts_ = 0, next_ = 0, finished_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

struct StageSeg {
  void *dst;          // device
  const void *src;    // pageable host
  size_t bytes;
};

class HostStager {
 public:
  ~HostStager();
  // Copies every segment to the device on stream s.  Returns when the host
  // memory has been read (the last DMAs may still run on s).  0 or a HIP
  // error code.
  hipError_t copy(hipStream_t s, const StageSeg *segs, int nseg);
  // Largest end index + 1 over the inclusive pairs' second entries, and the
  // key arena bytes of the packed keys (key pairs only), computed by the
  // workers.
  void extents(const int32_t *kidx, const int64_t *vidx, uint32_t n, uint64_t *kb, uint64_t *vb, uint64_t *ks,
               int max_nibbles);
  // Batches below this many bytes are copied directly (lower latency).
  static constexpr size_t kMinBytes = 8u << 20;

 private:
  hipError_t init();
  static constexpr int kBufs = 4;
  static constexpr size_t kChunk = 8u << 20;
  bool ready_ = false;
  void *pin_[kBufs] = {};
  hipEvent_t ev_[kBufs] = {};
  bool pending_[kBufs] = {};
  WorkerPool *pool_ = nullptr;
};

}  // namespace gmpt
