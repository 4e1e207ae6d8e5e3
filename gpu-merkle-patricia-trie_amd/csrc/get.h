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
// Splits a device batch by rw flags (0 read, 1 write) keeping batch order:
// read key pairs to rk, write key/value pairs (and host pointers, trie ids) to
// wk / wv / whp / wtid; *nread (device) receives the read count, and err a
// nonzero code for a flag other than 0 or 1.  isr / pos are n-entry workspace.
hipError_t rw_split(const uint8_t *flags, uint32_t n, const int32_t *kidx, const int64_t *vidx, const uint64_t *hps,
                    const int32_t *tid, uint32_t *isr, uint32_t *pos, uint32_t *nread, int32_t *rk, int32_t *rtid,
                    int32_t *wk, int64_t *wv, uint64_t *whp, int32_t *wtid, uint32_t *err, void *temp,
                    hipStream_t s);

// Exports every node on the paths of the (device-staged) keys once, as the
// reference's nodeset (libgmpt.h:13-18): 32-byte hash slots, RLP encodings,
// inclusive [start, end] u64 pairs.  Host arrays are new[]-allocated.
int export_paths(Forest *f, const BatchDesc &keys, uint8_t **hashes, uint8_t **encs, uint64_t **enc_indexs,
                 uint64_t *num);

}  // namespace gmpt
