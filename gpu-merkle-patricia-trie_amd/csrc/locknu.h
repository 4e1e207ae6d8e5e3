// locknu.h -- LockNU: latch-coupled concurrent insert (host launch interface).
#pragma once
#include "phasenu.h"

namespace gmpt {

// Inserts a.n keys (already packed by nu_prepare_keys) into their tries.
// Uses a.base_branch / base_short / base_value: per key i the slots
// base_branch + i, base_short + 2i (+1), base_value + i.
hipError_t lock_insert(const NuArgs &a, uint64_t key_cap, hipStream_t s);

}  // namespace gmpt
