// noalloc.hpp -- the fixed device scratch of the allocation-free entry point
// (stq_find_pairs_device_noalloc, noalloc.hip).
#pragma once

#include <cstddef>
#include <cstdint>

#include "../../include/stq.h"

namespace stqamd {

constexpr size_t kNoAllocScratchBytes = (size_t)STQ_NOALLOC_SCRATCH_BYTES;
constexpr int64_t kNoAllocMaxBoxes = STQ_NOALLOC_MAX_BOXES;

// Device address of the scratch array on `device`, which must be the current
// device.  Looked up once per device (hipGetSymbolAddress) and kept: each GPU
// holds its own copy of the array at its own address.  Returns a status.
int noalloc_scratch(int device, void **ptr);

}  // namespace stqamd
