// get.h -- point lookups and node export (host launch interface).
#pragma once
#include "common.h"
#include "engine.h"

namespace gmpt {

struct GetArgs {
  Pools P;
  const uint8_t *kh;
  const int32_t *kidx;
  const int32_t *tid;  // nullable
  uint32_t n;
  uint64_t kh_bytes;   // bytes in kh
  uint32_t ntries;
  uint64_t *hp;        // out: caller host pointer stored with the value (0 if none)
  uint64_t *addr;      // out: device address of the value bytes
  int32_t *size;       // out: value size, -1 when the key is absent
};

hipError_t launch_gets(const GetArgs &g, hipStream_t s);
hipError_t launch_gather_values(const uint64_t *addr, const int32_t *size, const uint64_t *off, uint32_t n,
                                uint8_t *out, hipStream_t s);
// Exports every node on the paths of the (device-staged) keys once, as the
// reference's nodeset (libgmpt.h:13-18): 32-byte hash slots, RLP encodings,
// inclusive [start, end] u64 pairs.  Host arrays are new[]-allocated.
int export_paths(Forest *f, const BatchDesc &keys, uint8_t **hashes, uint8_t **encs, uint64_t **enc_indexs,
                 uint64_t *num);

}  // namespace gmpt
