// Stable LSD radix sort, prefix scans and run-length encoding on the device
// (device_sort.hip), the sort and scan primitives of the action state, the
// scorer ranking, the patch bins and the classical reducer.
//
// Every entry point follows one calling convention: called with a null
// `temp` it stores the workspace bytes it needs in `temp_bytes` and queues
// nothing; called with a workspace of at least that size it queues its
// kernels on `stream` and returns. Inputs are never modified.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace blitz_hip {

// Sorts (key, value) pairs by bits [begin_bit, end_bit) of the key,
// ascending, stable (equal keys keep their input order). `values_in` /
// `values_out` may both be null (keys only).
hipError_t radix_sort_pairs(void* temp, size_t& temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                            const uint32_t* values_in, uint32_t* values_out, uint32_t n, int begin_bit, int end_bit,
                            hipStream_t stream);
hipError_t radix_sort_pairs(void* temp, size_t& temp_bytes, const uint64_t* keys_in, uint64_t* keys_out,
                            const uint32_t* values_in, uint32_t* values_out, uint32_t n, int begin_bit, int end_bit,
                            hipStream_t stream);
// Pairs by float key, descending, stable: NaN-free keys ordered as the
// floats are (-0 and +0 equal), ties in input order.
hipError_t radix_sort_pairs_descending(void* temp, size_t& temp_bytes, const float* keys_in, float* keys_out,
                                       const uint32_t* values_in, uint32_t* values_out, uint32_t n,
                                       hipStream_t stream);
// out[i] = in[0] + ... + in[i-1] (out[0] = 0); in and out may alias.
hipError_t exclusive_sum(void* temp, size_t& temp_bytes, const uint32_t* in, uint32_t* out, uint32_t n,
                         hipStream_t stream);
hipError_t exclusive_sum(void* temp, size_t& temp_bytes, const uint64_t* in, uint64_t* out, uint32_t n,
                         hipStream_t stream);
// out[i] = max(in[0], ..., in[i]); in and out may alias.
hipError_t inclusive_max(void* temp, size_t& temp_bytes, const uint32_t* in, uint32_t* out, uint32_t n,
                         hipStream_t stream);
// Runs of equal consecutive keys: unique[r], counts[r] for r < *runs.
hipError_t run_length_encode(void* temp, size_t& temp_bytes, const uint64_t* keys, uint64_t* unique, uint32_t* counts,
                             uint32_t* runs, uint32_t n, hipStream_t stream);

}  // namespace blitz_hip
