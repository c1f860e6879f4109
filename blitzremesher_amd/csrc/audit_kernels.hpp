// Launch wrappers for the coverage-audit kernels (audit_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>

#include "audit_types.hpp"

namespace blitz_hip {
void launch_fill(uint64_t* p, size_t n, uint64_t value, hipStream_t stream);
// tasks: device array of `count` tasks; grid covers max_faces per task.
void launch_raster(const RasterTask* tasks, uint32_t count, uint32_t max_faces, hipStream_t stream);
// pairs/summaries: device arrays of `count` views sharing one raster side.
void launch_compare(const ViewPair* pairs, ViewSummary* summaries, uint32_t count, uint32_t tiles,
                    hipStream_t stream);
void launch_distance(const ViewPair* pairs, ViewSummary* summaries, uint32_t count, uint32_t size, uint32_t tiles,
                     hipStream_t stream);
}  // namespace blitz_hip
