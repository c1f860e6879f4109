// Launch wrappers for the coverage-audit kernels (audit_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>

#include "audit_types.hpp"

namespace blitz_hip {
void launch_fill(uint64_t* p, size_t n, uint64_t value, hipStream_t stream);
// *differs |= any byte of a[0..bytes) != b[0..bytes) (the caller zeroes it).
void launch_differ(const void* a, const void* b, size_t bytes, uint64_t* differs, hipStream_t stream);
// tasks: device array of `count` projections; grid covers max_vertices each.
void launch_project(const ProjectTask* tasks, uint32_t count, uint32_t max_vertices, hipStream_t stream);
// tasks: device array of grid.tasks raster tasks sharing one raster side.
void launch_raster(const RasterTask* tasks, const RasterGrid& grid, uint32_t tiles, hipStream_t stream);
// Band and chunk split of `tasks` raster tasks of `tiles` x `tiles` words
// with up to `max_faces` faces each.
RasterGrid make_raster_grid(uint32_t tasks, uint32_t tiles, uint32_t max_faces);
// pairs/summaries: device arrays of `count` views sharing one raster side.
void launch_compare(const ViewPair* pairs, ViewSummary* summaries, uint32_t count, uint32_t tiles,
                    hipStream_t stream);
void launch_distance(const ViewPair* pairs, ViewSummary* summaries, uint32_t count, uint32_t size, uint32_t tiles,
                     hipStream_t stream);
// The directed Hausdorff pass alone (pairs without row-distance planes).
void launch_hausdorff(const ViewPair* pairs, ViewSummary* summaries, uint32_t count, uint32_t tiles,
                      hipStream_t stream);
// Face split of mesh a (reference) and mesh b (candidate) over one vertex set
// (face_split.hip). A mesh's triple table (face_table_slots(faces) slots of
// 16 bytes) is built once by launch_face_table. Scratch of a split: `match` of b_faces
// words, `matched` of a_faces bytes rounded up to 8, `block_counts` of
// 3 * face_split_blocks() words; `out` lists sized for b_faces shared,
// a_faces a-only and b_faces b-only entries.
uint32_t face_table_slots(uint32_t faces);
uint32_t face_split_blocks(uint32_t a_faces, uint32_t b_faces);
void launch_face_table(const uint32_t* idx, const uint8_t* two, uint32_t faces, uint4* table, hipStream_t stream);
void launch_face_split(const uint32_t* a_idx, const uint8_t* a_two, uint32_t a_faces, const uint4* a_table,
                       const uint32_t* b_idx, const uint8_t* b_two, uint32_t b_faces, uint32_t* match,
                       uint8_t* matched, uint32_t* block_counts, const FaceSplit& out, hipStream_t stream);
}  // namespace blitz_hip
