// server.h -- continuous-batching insert+hash service over HBM-resident tries.
//
// Requests (one host batch of key/value pairs for one trie) queue per lane; a
// lane owns a Forest (its own HIP stream) holding the tries t with
// t % lanes == lane.  Each lane's worker repeatedly takes every queued
// request it can admit -- at most one per trie (a trie's requests complete in
// arrival order, each with its own post-state root) and only while the
// lane's node-pool bytes plus the admitted requests' estimates stay within
// the lane's share of the memory budget -- inserts them as one multi-trie
// batch, hashes once, and completes each request with its trie's root.
// Requests that arrive while a batch is on the GPU join the next batch;
// lanes run concurrently, so batches of different lanes overlap on the device.
#pragma once
#include <stdint.h>

#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "engine.h"

namespace gmpt {

struct SrvStats {
  uint64_t submitted = 0, completed = 0, failed = 0;
  uint64_t batches = 0, batched_requests = 0, batched_keys = 0;
  uint64_t deferred_mem = 0;  // times a request waited because the memory budget was full
  double busy_s = 0;          // summed lane time spent in insert + hash
};

class Server {
 public:
  // mode 0: PhaseNU inserts, 1: LockNU inserts.  mem_budget = 0 takes 80% of
  // the device's free memory at creation.
  Server(int device, uint32_t ntries, uint32_t lanes, uint64_t mem_budget, uint32_t max_batch_keys, int mode);
  ~Server();
  bool ok() const { return ok_; }
  int submit(uint32_t trie, const BatchDesc &host_batch, uint64_t *id);
  // state: 0 queued, 1 running, 2 done, -1 failed (root, latency valid when 2)
  int poll(uint64_t id, int *state, uint8_t *root, int *root_size, double *latency_ms);
  int wait(uint64_t id, int *state, uint8_t *root, int *root_size, double *latency_ms);
  int drain();
  SrvStats stats();
  // latencies (ms, submit -> root available) of requests completed so far, in id order
  std::vector<double> latencies();

 private:
  struct Req {
    uint32_t trie = 0, n = 0;
    std::vector<uint8_t> keys;
    std::vector<int32_t> kidx;
    std::vector<uint8_t> vals;
    std::vector<int64_t> vidx;
    uint64_t need = 0;
    double t_submit = 0, t_done = 0;
    int state = 0;
    uint8_t root[32] = {0};
    int root_size = 0;
    std::string err;
  };
  struct Pinned {  // page-locked host staging buffer (grown on demand)
    void *p = nullptr;
    size_t cap = 0;
    int reserve(size_t bytes);
    ~Pinned();
  };
  struct Lane {
    std::unique_ptr<Forest> f;
    std::deque<uint64_t> q;
    std::thread th;
    uint64_t budget = 0;
    uint32_t local_tries = 0;
    Pinned keys, kidx, vals, vidx, tid;  // the concatenated batch
    std::vector<uint8_t> roots;
    std::vector<int32_t> root_sizes;
  };
  void worker(uint32_t lane);
  int run_batch(Lane &L, const std::vector<Req *> &batch);

  int device_ = 0;
  uint32_t ntries_ = 0, nlanes_ = 0, max_batch_keys_ = 0;
  int mode_ = 0;
  bool ok_ = false, stop_ = false;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::vector<std::unique_ptr<Req>> reqs_;
  std::vector<std::unique_ptr<Lane>> lanes_;
  SrvStats st_;
};

}  // namespace gmpt
