// Entry points that allocate nothing (noalloc.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace blitz_hip {

// Largest raster side of coverage_distance_noalloc: its fixed array holds two
// tile-major masks of 8192 x 8192 samples (16 MiB) per device.
constexpr uint32_t kNoAllocMaxSide = 8192;

// The coverage distance of two row-major packed masks (bit y*size + x) in
// device memory, written to the device double `out`; queued on `stream`,
// nothing allocated.
void coverage_distance_noalloc(int device, const uint64_t* a, const uint64_t* b, uint32_t size, uint8_t supersample,
                               double* out, hipStream_t stream);

}  // namespace blitz_hip
