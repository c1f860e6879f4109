// Plain data shared by the host controller and the audit kernels.
#pragma once
#include <cstdint>

namespace blitz_hip {

struct F3 {
    float x, y, z;
};

// Same layout as blitz::Camera (reference include/blitz/evaluate.hpp:46).
struct Camera {
    F3 right, up, forward;
    double distance, focal, scale;
    uint8_t perspective;
};

struct Bounds {
    F3 center;
    double radius;
};

// Coverage masks are stored tile-major: one 64-bit word per 8x8 sample tile,
// word = ty*tiles + tx, bit = (y&7)*8 + (x&7). A 64-lane wavefront owns one
// tile with lane == bit, so a ballot over the tile is the tile's word.
constexpr uint32_t kTile = 8;

// Projection of one vertex set from one camera at one supersampling.
struct ProjectTask {
    const float4* positions;  // xyz + pad, one per vertex
    double2* xy;              // projected x, y per vertex (FP64, reference order)
    double* z;                // view depth per vertex; null for orthographic views
    Camera camera;
    F3 center;
    uint32_t vertices;
    uint32_t size;            // raster side in samples
    uint8_t ss;
};

// One raster job: one mesh's faces over one projected vertex set.
struct RasterTask {
    const double2* xy;
    const double* z;           // null for orthographic views
    const uint32_t* indices;   // 3 per face
    const uint8_t* two_sided;  // per-face flag or null (all one-sided)
    uint64_t* mask;            // tile-major words, tiles*tiles
    uint32_t* status;          // [0] first out-of-frame face, [1] behind-camera flag
    uint32_t faces;            // faces rasterized: [0, faces)
    uint32_t size;             // raster side in samples
    uint32_t tiles;            // ceil(size / 8)
    uint8_t force_two;
    // Face lists of a split reference/candidate pair (face_split.hip): entry
    // f's id in its own mesh (bit 31: two-sided), which replaces f in the
    // status words and `two_sided`; null for a whole mesh.
    const uint32_t* ids = nullptr;
    // Shared faces: the second mesh's mask, status words and ids.
    uint64_t* mask2 = nullptr;
    uint32_t* status2 = nullptr;
    const uint32_t* ids2 = nullptr;
    // Device-side face count (list length known only on the device), at
    // most `faces`; null: `faces`.
    const uint32_t* count = nullptr;
    // Flush with atomicOr even when the task has one chunk (another task
    // writes the same mask).
    uint8_t or_flush = 0;
};

// Shared / reference-only / candidate-only face lists of a split pair:
// triples (3 per face) and ids (bit 31: two-sided); counts[0..2] = lengths.
struct FaceSplit {
    uint32_t* shared_tri;
    uint32_t* shared_ids_a;
    uint32_t* shared_ids_b;
    uint32_t* a_tri;
    uint32_t* a_ids;
    uint32_t* b_tri;
    uint32_t* b_ids;
    uint32_t* counts;
};

// Launch shape of the band raster: every task is split into `bands` bands of
// `band_rows` tile rows (one LDS-resident band per workgroup) and its faces
// into `chunks` contiguous ranges.
struct RasterGrid {
    uint32_t tasks, bands, chunks, band_rows;
};
constexpr uint32_t kRasterThreads = 1024;
constexpr uint32_t kBandWords = 16384;  // 128 KiB of LDS per workgroup

// Per-view comparison of a reference mask A and a candidate mask C.
struct ViewPair {
    const uint64_t* a;
    const uint64_t* c;
    uint16_t* dist_a;  // row distances to the nearest A sample (queries from C)
    uint16_t* dist_c;  // row distances to the nearest C sample (queries from A)
    uint32_t size, tiles;
    // Verdict-only audits: a squared distance at or above this settles the
    // view as failing, so the Hausdorff pass may stop there with a lower
    // bound. 0xFFFFFFFF: exact maxima.
    uint32_t fail_sq = 0xFFFFFFFFu;
};

// Written by the compare and distance kernels, read back by the host.
struct ViewSummary {
    unsigned long long count_a, count_c, changed, total;
    uint32_t max_ac;  // max over A\C of the squared distance to C (pre-rounding)
    uint32_t max_ca;  // max over C\A of the squared distance to A
    uint32_t pad[2];
};

constexpr uint16_t kNoSite = 0xFFFF;

}  // namespace blitz_hip
