// locknu.h -- LockNU: latch-coupled concurrent insert (host launch interface).
#pragma once
#include "phasenu.h"

namespace gmpt {

// Inserts the batch's unique keys (packed by nu_prepare_keys, sorted and
// deduplicated) into their tries.  The k-th new key (in sorted order) takes
// the slots base_branch + k, base_short + 2k (+1), base_value + k; keys that
// exist are rewritten in place.  temp: scan workspace.
hipError_t lock_insert(const NuArgs &a, uint64_t key_cap, void *temp, hipStream_t s);

}  // namespace gmpt
