// noalloc.hpp -- the scratch buffer of the entry point that allocates nothing.
#pragma once

#include <cstddef>
#include <cstdint>

namespace pscbfm {

// Size of the per-device scratch array: 32 MiB, the copy-back of up to
// kNoAllocMaxMonomers monomers (12 bytes each: x, y, z as int32).
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
constexpr uint32_t kNoAllocMaxMonomers = uint32_t(kNoAllocScratchBytes / (3 * sizeof(int32_t)));

// Device address of the scratch array on `device` (looked up with the device
// current on the first call for that device, then kept).  Allocates nothing.
void *noalloc_scratch(int device);

}  // namespace pscbfm
