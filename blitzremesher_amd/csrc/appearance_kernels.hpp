// Kernels of the normals and attributes audit profiles
// (appearance_kernels.hip): the attribute raster, the normal angle and the
// attributed (product-metric) distance of src/evaluate.cpp:83-95,125-270.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "audit_types.hpp"

namespace blitz_hip {

// One attribute raster (or overlap count): one mesh's faces over one
// projected vertex set, drawn from per-tile face bins. Outputs are tile-major like the masks
// (sample = tile * 64 + (y & 7) * 8 + (x & 7)), every sample written.
struct AttrTask {
    RasterTask raster;                      // faces [0, faces), projection, sidedness, side, tiles
    const double* depth_z;                  // view depth per vertex (every view)
    const float4* positions;                // xyz per vertex
    const float4* normals;                  // unit vertex normals (xyz) or null
    const float4* colors;                   // rgba / 255 per vertex or null
    const uint16_t* materials;              // per face or null
    const unsigned long long* bin_offsets;  // this task's tiles*tiles cells (+1 past the end)
    const uint32_t* bin_faces;
    float4* normal_out;  // xyz; w = 1 when visible, 0 otherwise
    float4* color_out;   // rgb; w = material id bits (uint32); null when not needed
    float* depth_out;    // optional: the stored float depth, +inf where nothing is visible
    double diameter;
    uint8_t perspective;
};

// One view's pair of attribute rasters (reference A, candidate C) and its
// device outputs: [0] max metric_acos of the normal angle over samples
// visible in both, [1] max squared attributed cost over both directions.
// Both are non-negative doubles reduced as their bit patterns.
struct AttrPair {
    const float4* normal_a;
    const float4* color_a;
    const float4* normal_c;
    const float4* color_c;    // colours may be null unless the profile is attributes
    unsigned long long* out;  // 2 per pair
    uint32_t size, tiles;
};

// Weights of src/evaluate.cpp:83-95.
struct CostSettings {
    double normal_weight, color_weight, material_weight;
    double limit;
    uint32_t profile;
    uint32_t ss;
};

void launch_attribute_raster(const AttrTask* tasks, uint32_t count, uint32_t tiles, hipStream_t stream);
// Per task: sums[2j] = overlap samples, sums[2j+1] = samples with any (zeroed
// by the caller).
void launch_overlap(const AttrTask* tasks, uint32_t count, uint32_t tiles, unsigned long long* sums,
                    hipStream_t stream);
void launch_normal_angle(const AttrPair* pairs, uint32_t count, uint32_t tiles, hipStream_t stream);
void launch_attributed(const AttrPair* pairs, uint32_t count, uint32_t tiles, const CostSettings& s,
                       hipStream_t stream);

}  // namespace blitz_hip
