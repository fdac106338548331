// Stable radix sort of (u32 key, u32 value) pairs (rl_sort.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace rl {

// Device scratch bytes launch_sort_pairs needs for n pairs.
size_t sort_pairs_temp_bytes(uint32_t n);

// Sorts n pairs by the low key_bits bits of the key, stably (equal keys keep their input order),
// ping-ponging between (keys, vals) and (keys_alt, vals_alt); *keys_out / *vals_out receive the
// buffers that hold the result. Keys must be below 2^key_bits.
hipError_t launch_sort_pairs(hipStream_t st, uint32_t* keys, uint32_t* keys_alt, uint32_t* vals, uint32_t* vals_alt,
                             uint32_t n, int key_bits, void* temp, size_t temp_bytes, uint32_t** keys_out,
                             uint32_t** vals_out);

}  // namespace rl
