// Host-side interface of rl_noalloc.hip: segment tracing through a per-device static scratch
// array instead of per-call device allocations.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "../../include/relight.h"
#include "rl_device.h"

namespace rl {

// Largest segment count of one call, and the scratch array's size in bytes (the device memory
// each GPU holds for it until the process ends): 256-B header, segment triples in, packed float4
// segments, raw hit records, rl_hit records and hit flags for kNoAllocMaxSegments segments.
constexpr uint32_t kNoAllocMaxSegments = 65536;
constexpr size_t kNoAllocScratchBytes = 6357248;

// Looks up the device address of the scratch array on `device` (must be the current device) and
// keeps it for later calls; idempotent per device.
hipError_t noalloc_setup(int device);

// rl_trace_segments on `device` with every device buffer inside the scratch array: n <=
// kNoAllocMaxSegments host segments in, host flags (any) or hit records (closest) out, rays cast
// in *out_rays. Returns after the work has finished. hipErrorNotInitialized if noalloc_setup has
// not run on the device, hipErrorInvalidValue if n is too large.
hipError_t noalloc_trace(int device, hipStream_t st, const SceneDev& S, int cap, uint32_t max_blocks, bool any,
                         const float* starts, const float* ends, uint32_t n, uint8_t* out_flags, rl_hit* out_hits,
                         unsigned long long* out_rays);

}  // namespace rl
