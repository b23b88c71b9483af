// noalloc.hpp -- the fixed per-device workspace behind the entry points that
// allocate nothing (noalloc.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>

namespace ige {

// Size of the workspace each device holds from the library's load until the
// process ends.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;

// Looks up the workspace's device address on `device`, which must be the
// current device; once per device (later calls return at once).
hipError_t noalloc_prepare(int device);
// The workspace's device address on `device` (NULL before noalloc_prepare).
void* noalloc_region(int device);
// Held by a call for as long as it uses the device's workspace, so calls from
// different host threads on one device take turns.
std::mutex& noalloc_mutex(int device);

}  // namespace ige
