// Kernels of the bound-reference patch audit (patch_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "audit_types.hpp"

namespace blitz_hip {

// Largest region (removed faces) and re-meshed region (added faces) one patch
// may have on the device path, and the largest dirty rectangle of tiles.
constexpr uint32_t kPatchMaxFaces = 512;
constexpr uint32_t kPatchRectWords = 1024;

struct PatchSummary {
    unsigned long long count_c, changed, total;
    uint32_t max_ac, max_ca;
    uint32_t flags;  // bit 0: the patch must take the full audit path in this view
    uint32_t pad;
};

struct PatchInput {
    const RasterTask* views;               // per view: reference faces over that view's projection
    const uint64_t* masks;                 // per view: reference mask, tiles*tiles words
    const unsigned long long* count_a;     // per view: samples covered by the reference
    const unsigned long long* bin_offsets; // views*tiles*tiles + 1
    const uint32_t* bin_faces;             // reference faces per (view, tile)
    const uint32_t* removed_offsets;       // patches + 1
    const uint32_t* removed;               // ascending face ids per patch
    const uint32_t* added_offsets;         // patches + 1 (faces)
    const uint32_t* added;                 // 3 vertex ids per added face
    PatchSummary* out;                     // views * patches, view-major
    uint32_t view_count, patches, tiles;
    uint8_t added_two_sided;  // added faces are drawn two-sided (material 0 of the reference)
};

// Per (view, face): number of 8x8 tiles of each drawable face's sample box.
void launch_bin_count(const RasterTask* views, uint32_t count, uint32_t max_faces, unsigned long long* counts,
                      hipStream_t stream);
// Per (view, face): face ids into their tiles' bins (cursor = bin offsets).
void launch_bin_fill(const RasterTask* views, uint32_t count, uint32_t max_faces, unsigned long long* cursor,
                     uint32_t* faces, hipStream_t stream);
size_t scan_workspace(size_t n);
void launch_exclusive_scan(void* temp, size_t temp_bytes, const unsigned long long* in, unsigned long long* out,
                           size_t n, hipStream_t stream);
void launch_patches(const PatchInput& in, hipStream_t stream);

}  // namespace blitz_hip
