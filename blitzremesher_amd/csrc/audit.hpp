// Host controller of the HIP coverage audit: the counterpart of the
// reference's AuditCuda (src/neural/internal.hpp:131-175) with the CPU
// audit's semantics (src/evaluate.cpp:312-352).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "audit_types.hpp"
#include "autotune.hpp"

namespace blitz_hip {

// Raised for settings the HIP backend does not implement (maps to
// BLITZ_UNAVAILABLE, as blitz::NeuralUnavailable does).
struct Unavailable : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what);

struct Settings {
    uint8_t profile = 1;  // EvalSettings defaults: Normals
    uint8_t supersample = 8, max_supersample = 32;
    bool force_two_sided = false, conservative_screen = false;
    uint16_t orthographic = 642, perspective = 64;
    uint32_t rotation_seed = 0xB1172026u;
    double screen_size = 512, limit = 2, max_changed_area = 1;
    double normal_weight = 180.0 / (10 * 3.14159265358979323846), color_weight = 4, material_weight = 4;
    std::function<bool()> cancelled;
    // Only `passed` is wanted (an executor's gate): the coverage distance of
    // a view may stop at a lower bound once it settles the view as failing.
    // Every other Measurement field of a failing view is then a bound.
    bool verdict_only = false;
};

struct Measurement {
    double error = 0, coverage = 0, coverage_upper = 0, changed_area = 0, normal_degrees = 0;
    uint32_t worst_view = 0, changed_area_worst_view = 0, views_evaluated = 0;
    uint8_t supersample = 0;
    bool complete = true, passed = true, resource_limited = false, cancelled = false;
};

// Host copy of a borrowed mesh, validated (reference src/mesh.cpp:92-106).
struct HostMesh {
    std::vector<float> positions;  // xyz
    std::vector<uint32_t> indices;
    std::vector<uint16_t> materials;
    std::vector<uint8_t> double_sided;
    std::vector<float> normals;    // xyz per vertex as given, or empty
    std::vector<uint8_t> colors;   // RGBA8 per vertex, or empty
    uint64_t digest[2] = {0, 0};      // streams compared by identical(): positions, normals, colors,
                                      // indices, materials, sidedness
    uint64_t pos_digest[2] = {0, 0};  // positions only: equal digests share one projection per view
    uint64_t vertex_digest[2] = {0, 0};  // digest state after the vertex streams (positions, normals, colors)
};

class Audit;

struct DeviceMesh {
    Audit* owner = nullptr;
    float4* positions = nullptr;
    uint32_t* indices = nullptr;
    uint8_t* two_sided = nullptr;  // per face; null when no face is two-sided
    // Open-addressing table of faces keyed by the ordered vertex triple
    // (face_split.hip), face_table_slots(faces) slots; null when not built.
    uint4* face_table = nullptr;
    // Appearance streams (include/blitz/mesh.hpp:20): unit vertex normals,
    // vertex colours / 255 and per-face materials; null when absent.
    float4* normals = nullptr;
    float4* colors = nullptr;
    uint16_t* materials = nullptr;
    uint32_t vertices = 0, faces = 0;
    uint64_t digest[2] = {0, 0};
    uint64_t pos_digest[2] = {0, 0};
    uint64_t vertex_digest[2] = {0, 0};
    uint64_t serial = 0;
    // Serial of the device mesh that owns the position buffer (this mesh's
    // own serial unless derived).
    uint64_t positions_serial = 0;
    // Host copies of the raw streams that identity compares and the device
    // does not hold as given: normals and colours (the device holds them
    // normalised / scaled) and the per-material sidedness table.
    std::shared_ptr<const std::vector<float>> raw_normals;
    std::shared_ptr<const std::vector<uint8_t>> raw_colors;
    std::vector<uint8_t> raw_double_sided;
    // Vertex streams (positions, normals, colours) owned by another device
    // mesh of the same audit (derive()): not freed by this one.
    bool borrowed_vertices = false;
    ~DeviceMesh();
};

Bounds bounds_of(const float* xyz, size_t count);
std::vector<Camera> cameras(const Bounds& b, double screen, uint16_t ortho, uint16_t persp, uint32_t seed);
void validate_settings(const Bounds& b, const Settings& s);
uint32_t raster_side(double screen, unsigned ss);  // ceil(screen + 8) * ss
// One view's coverage-profile result (src/evaluate.cpp:271-298) and the
// refinement stop rule (src/evaluate.cpp:336-337).
Measurement finish_view(double coverage, uint64_t changed, uint64_t total, bool clipped, const Settings& s,
                        uint8_t ss);
// The same view under the normals and attributes profiles: normal_acos is
// the largest metric_acos of the normal angle over samples visible in both
// rasters, attributed the attributed distance (used only when the coverage
// bound is within the limit) and previous_degrees the view's normal angle
// from its earlier refinement passes.
Measurement finish_view_attributed(double coverage, uint64_t changed, uint64_t total, bool clipped,
                                   double normal_acos, double attributed, double previous_degrees,
                                   const Settings& s, uint8_t ss);
// Whether the attributed distance can be non-zero (src/evaluate.cpp:218-219).
bool attributed_weights(const Settings& s);
bool settled(const Measurement& cur, const Settings& s, unsigned ss);

struct Timing {
    uint32_t raster_tasks = 0, distance_views = 0, passes = 0;
    double raster_ms = 0, compare_ms = 0, distance_ms = 0, total_ms = 0;
    uint64_t raster_triangles = 0, mask_bytes = 0;
    // Device time of the projection and band-raster launches alone (events
    // immediately around each launch).
    double project_ms = 0, band_ms = 0;
    uint32_t band_launches = 0;
    double patch_ms = 0;      // patch kernel (device time)
    uint32_t patch_fallbacks = 0;
    double appearance_ms = 0;  // attribute rasters, normal angle and attributed distance (device time)
    uint32_t attributed_views = 0;
    uint32_t split_launches = 0;  // raster launches drawn from a reference/candidate face split
    double split_ms = 0;          // face split kernels (device time)
    uint32_t raster_chunks = 0;   // face chunks per task of the last raster launch
    uint32_t tuned_launches = 0;  // raster launches whose shape was measured in this call
};

// A reference bound for patch audits (patch_audit.hpp).
struct Binding;
// Patches as CSR arrays: removed source face ids (ascending) and added faces
// (3 vertex ids each) of every patch.
struct PatchSet {
    const uint32_t* removed_offsets = nullptr;  // count + 1
    const uint32_t* removed = nullptr;
    const uint32_t* added_offsets = nullptr;    // count + 1, in faces
    const uint32_t* added = nullptr;
    size_t count = 0;
};

class Audit {
  public:
    // cache_rasters: reuse the reference's coverage masks across audits
    // against the same uploaded reference (the reference's RasterMemo,
    // src/neural/evaluate.cu:386-511); results are unchanged.
    // face_split: 0 automatic, 1 never, 2 whenever a reference/candidate
    // pair shares a vertex set (raster_launch).
    // autotune: pick the raster launch shape by timing (autotune.hip).
    Audit(int device, uint64_t memory_bytes, bool cache_rasters = false, uint8_t face_split = 0,
          bool autotune = true);
    ~Audit();
    Audit(const Audit&) = delete;
    Audit& operator=(const Audit&) = delete;

    std::unique_ptr<DeviceMesh> upload(const HostMesh& m);
    // A mesh with `base`'s vertex streams (borrowed: `base` must outlive it)
    // and its own triangles: `indices` (3 per face, each < base.vertices),
    // per-face `materials` (empty or one per face) and the per-material
    // sidedness table `double_sided`. `digest` is the identity digest of the
    // whole mesh, as host_mesh computes it.
    std::unique_ptr<DeviceMesh> derive(const DeviceMesh& base, const std::vector<uint32_t>& indices,
                                       const std::vector<uint16_t>& materials,
                                       const std::vector<uint8_t>& double_sided, const uint64_t digest[2]);
    Measurement evaluate(const DeviceMesh& reference, const DeviceMesh& candidate, const Bounds& b,
                         const Settings& s);
    std::vector<Measurement> evaluate_batch(const DeviceMesh& reference,
                                            const std::vector<const DeviceMesh*>& candidates, const Bounds& b,
                                            const Settings& s);
    // Row-major packed mask of one view; returns false when over the sample cap.
    bool raster_row_major(const DeviceMesh& m, const Bounds& b, const Camera& c, double screen, uint8_t ss,
                          bool force_two, std::vector<uint64_t>& words, uint32_t& size, bool& clipped);
    double coverage_distance(const std::vector<uint64_t>& a, const std::vector<uint64_t>& b, uint32_t size,
                             uint8_t ss);
    // Attribute raster of one view (src/evaluate.cpp:125-189), row-major:
    // covered, visible, stored depth, normal xyz, colour rgba, material.
    // Returns false when over the sample cap.
    struct PixelRaster {
        uint32_t size = 0;
        bool clipped = false;
        std::vector<uint8_t> covered, visible;
        std::vector<float> depth, normal, color;
        std::vector<uint16_t> material;
    };
    bool raster_pixels(const DeviceMesh& m, const Bounds& b, const Camera& c, double screen, uint8_t ss,
                       bool force_two, PixelRaster& out);
    // Mean over views of the overdraw samples per covered pixel under the
    // top-left fill rule (overlap_cuda, src/neural/evaluate.cu:2665-2685);
    // infinite when any view is clipped.
    double overlap(const DeviceMesh& m, const Bounds& b, double screen, uint16_t ortho, uint16_t persp,
                   uint32_t seed);
    // Binds a reference: per view its projection, mask, covered count and
    // per-tile face bins at s.supersample (src/neural/internal.hpp:141-144).
    std::unique_ptr<Binding> bind(const HostMesh& reference, const Bounds& b, const Settings& s);
    // out[i] equals evaluate(reference, candidate_i), candidate_i being the
    // reference with patch i applied; `fallbacks` receives how many patches
    // took the full audit path.
    std::vector<Measurement> evaluate_patches(Binding& binding, const PatchSet& patches, const Settings& s,
                                              uint32_t* fallbacks);
    // blitz::identical (src/evaluate.cpp:71-82): positions, normals, colours,
    // indices, materials and the sidedness table byte-equal. Digests decide
    // only when they differ; equal digests are confirmed stream by stream.
    bool identical(const DeviceMesh& a, const DeviceMesh& b);
    // Position streams byte-equal (same owner, or compared once per pair of
    // owners and remembered).
    bool same_positions(const DeviceMesh& a, const DeviceMesh& b);
    const Timing& last_timing() const { return timing_; }
    int device() const { return device_; }
    uint64_t memory_bytes() const { return memory_bytes_; }
    bool cache_rasters() const { return cache_rasters_; }
    hipStream_t stream() const { return stream_; }
    void release(DeviceMesh* m);

  private:
    struct Item {
        uint32_t candidate;  // index into the candidate list
        uint32_t view;
        uint8_t ss;
    };
    struct ViewResult {
        double coverage;
        uint64_t changed, total;
        bool clipped;
        double normal_acos;  // appearance profiles: largest normal angle (metric_acos)
        double attributed;   // appearance profiles: attributed distance, 0 when not needed
    };
    struct RasterJob {
        const DeviceMesh* mesh;
        uint32_t view;
        uint64_t* mask;  // tile-major, zeroed by the caller
    };
    // Rasterizes every job at one supersampling; status receives per job the
    // first out-of-frame face (0xFFFFFFFF if none) and the behind-camera flag.
    // With `drawn` set, also returns every job's raster task as drawn (faces
    // cut at the first out-of-frame face) and `depth` its vertex set's view
    // depths, computed for every view.
    void rasterize(const std::vector<RasterJob>& jobs, const std::vector<Camera>& cams, const Bounds& b, uint8_t ss,
                   uint32_t size, bool force_two, std::vector<uint32_t>& status,
                   std::vector<RasterTask>* drawn = nullptr, std::vector<const double*>* depth = nullptr);
    // rasterize() in two halves. raster_launch queues the projections, the
    // band raster and the copy of the status words into pinned host memory,
    // and waits for nothing; raster_finish, called once the stream has been
    // synchronised, reads the status words, redraws the masks cut at an
    // out-of-frame face (launch and wait) and says whether there were any.
    struct RasterLaunch {
        std::vector<RasterTask> tasks;
        std::vector<const double*> depth;  // per job: its vertex set's view depths, when requested
        bool split = false;                // drawn from a reference/candidate face split
        size_t words = 0;
        uint32_t tiles = 0;
    };
    // `pairs` (reference job, candidate job) of one view each, all with the
    // same two meshes: drawn from their face split when they share a vertex
    // set and the raster is large enough to repay it.
    void raster_launch(const std::vector<RasterJob>& jobs, const std::vector<Camera>& cams, const Bounds& b,
                       uint8_t ss, uint32_t size, bool force_two, bool want_depth, RasterLaunch& out,
                       const std::vector<std::pair<uint32_t, uint32_t>>* pairs = nullptr);
    // Queues the face split of reference a and candidate b (face_split.hip)
    // into scratch slots 40-47.
    FaceSplit split_faces(const DeviceMesh& a, const DeviceMesh& b);
    bool raster_finish(RasterLaunch& l, std::vector<uint32_t>& status, std::vector<RasterTask>* drawn = nullptr,
                       std::vector<const double*>* depth = nullptr);
    void launch_tasks(const std::vector<RasterTask>& list, size_t slot, uint32_t tiles);
    // Per (job, tile) bins of the faces each drawn job draws, face ids into
    // `faces`; job j's cells start at offsets + j*tiles*tiles.
    void build_bins(const std::vector<RasterTask>& drawn, const unsigned long long*& offsets,
                    const uint32_t*& faces);
    // Attribute rasters of drawn jobs: job j writes words*64 samples at
    // normal_out/color_out (and depth_out, when given) + j*words*64.
    void attribute_rasters(const std::vector<RasterJob>& jobs, const std::vector<RasterTask>& drawn,
                           const std::vector<const double*>& depth, const std::vector<Camera>& cams, const Bounds& b,
                           float4* normal_out, float4* color_out, float* depth_out);
    void run_items(const DeviceMesh& reference, const std::vector<const DeviceMesh*>& candidates,
                   const std::vector<Camera>& cams, const Bounds& b, const Settings& s, const std::vector<Item>& items,
                   std::vector<ViewResult>& out);
    void* scratch(size_t slot, size_t bytes);
    template <class T> T* scratch_as(size_t slot, size_t count) {
        return static_cast<T*>(scratch(slot, count * sizeof(T)));
    }

    // Reference masks kept across audits, per (supersampling, view), valid
    // for one (reference upload, bounds, screen, view set, sidedness).
    struct CachedMask {
        uint64_t* mask;
        bool clipped;
    };
    struct RasterCache {
        uint64_t serial = 0;
        Bounds bounds{};
        double screen = 0;
        uint16_t ortho = 0, persp = 0;
        uint32_t seed = 0;
        bool two = false;
        size_t bytes = 0;
        std::map<std::pair<uint8_t, uint32_t>, CachedMask> masks;
    };
    void cache_context(const DeviceMesh& reference, const Bounds& b, const Settings& s);
    void cache_clear();

    int device_;
    uint64_t memory_bytes_;
    bool cache_rasters_ = false;
    uint8_t face_split_ = 0;
    RasterTuner tuner_;
    uint32_t* pinned_status_ = nullptr;  // status words of the last raster launch (pinned host memory)
    size_t pinned_status_words_ = 0;
    RasterCache cache_;
    std::map<std::pair<uint64_t, uint64_t>, bool> position_pairs_;  // owner serials -> equal
    bool device_equal(const void* a, const void* b, size_t bytes);
    hipStream_t stream_{};
    hipEvent_t events_[11]{};
    std::vector<std::pair<void*, size_t>> slots_;
    Timing timing_;
};

// Restores the caller's current device on scope exit (reference
// src/neural/cuda.cuh:32-45).
struct DeviceGuard {
    int previous = -1;
    explicit DeviceGuard(int device);
    ~DeviceGuard();
};

}  // namespace blitz_hip
