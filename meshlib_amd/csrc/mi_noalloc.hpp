// mi_noalloc.hpp -- fixed device scratch of the entry point that allocates nothing (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>

namespace mi {

// Size of the per-GPU scratch array: the Morton sort (four n-word arrays and the radix sort's
// temporary storage) and the projection walk's split scratch (about 9 bytes per point) for up to
// kNoAllocMaxPoints points.
constexpr size_t kNoAllocScratchBytes = size_t(96) << 20;
constexpr long long kNoAllocMaxPoints = 1LL << 21;

// device address of `device`'s copy of the scratch array (looked up once per device, with that
// device current; the caller's current device is restored)
hipError_t noAllocScratch(int device, void** ptr);

} // namespace mi
