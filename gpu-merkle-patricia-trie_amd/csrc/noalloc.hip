// noalloc.hip -- gmpt_build_root_noalloc: the root of the trie one batch
// builds from empty, computed without any allocation call.
//
// Every device buffer of the call (node pools, key arena, sort and hashing
// workspace, staging of host batches) is carved from `scratch`, a fixed
// device array that each GPU holds for the life of the process.  The forest
// object and its host mirrors live on the caller's stack.  Host code never
// names `scratch` itself for device memory: its device address is looked up
// once per device (hipGetSymbolAddress) and that address is what the kernels
// get.
#include <string.h>

#include <mutex>
#include <string>

#include "../../include/gmpt_amd.h"
#include "engine.h"

namespace {

constexpr size_t kScratchBytes = (size_t)GMPT_NOALLOC_SCRATCH_MIB << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_mu[kMaxDevices];        // one call at a time per device: they share its region
uint8_t *g_ptr[kMaxDevices] = {};    // scratch's device address on each device (set under g_mu)

int fail(const char *msg, int code = -3) {
  gmpt::set_error(msg);
  return code;
}

}  // namespace

extern "C" int gmpt_build_root_noalloc(const gmpt_batch *b, int mode, uint8_t out[32], int *size) {
  if (!b || !out || !size) return fail("null argument");
  *size = 0;
  if (mode != 0 && mode != 1) return fail("mode must be 0 (PhaseNU) or 1 (LockNU)");
  if (b->n < 0) return fail("negative batch size");
  if (b->n > GMPT_NOALLOC_MAX_KEYS) {
    gmpt::set_error("batch of " + std::to_string(b->n) + " keys is above the no-allocation entry point's limit of " +
                    std::to_string(GMPT_NOALLOC_MAX_KEYS) + " keys");
    return -3;
  }
  if (b->n > 0 && (!b->keys_hexs || !b->keys_indexs || !b->values_bytes || !b->values_indexs))
    return fail("batch without keys or values");
  if (b->n > 0 && b->on_device && (b->keys_bytes <= 0 || b->values_bytes_total < 0))
    return fail("device batch needs keys_bytes and values_bytes_total");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return fail("no usable current device", -1);
  std::lock_guard<std::mutex> lock(g_mu[dev]);
  if (!g_ptr[dev]) {
    // once per device, with that device current: the array's address there
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      gmpt::set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return -1;
    }
    g_ptr[dev] = (uint8_t *)p;
  }
  if (b->n == 0) return 0;  // the empty trie has no root
  gmpt::BatchDesc d;
  d.keys_hexs = b->keys_hexs;
  d.keys_indexs = (const int32_t *)b->keys_indexs;
  d.values_bytes = b->values_bytes;
  d.values_indexs = b->values_indexs;
  d.values_hps = (const uint64_t *)b->values_hps;
  d.trie_ids = nullptr;  // one trie
  d.n = (uint32_t)b->n;
  d.keys_bytes = (uint64_t)(b->keys_bytes > 0 ? b->keys_bytes : 0);
  d.values_total = (uint64_t)(b->values_bytes_total > 0 ? b->values_bytes_total : 0);
  d.on_device = b->on_device != 0;
  // the trie lives for this call only, so it may point into the caller's
  // device values instead of copying them
  d.adopt_values = d.on_device;
  gmpt::Arena arena;
  arena.base = g_ptr[dev];
  arena.cap = kScratchBytes;
  gmpt::HostMirror host;
  memset(&host, 0, sizeof(host));
  gmpt::Forest f(dev, 1, &arena, &host, nullptr);
  if (!f.ok()) return -1;
  void *targets = nullptr;
  int nt = 0;
  int r = mode == 0 ? f.puts_2phase(d, &targets, &nt) : f.puts_latching(d, &targets, &nt);
  if (r) return r;
  if ((r = f.hash(targets, nt))) return r;
  return f.root_hash(0, out, size);
}
