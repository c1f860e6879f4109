/* MI355X (gfx950) HIP backend for BlitzRemesher's audit + scorer hot path.
 *
 * This is the thin C ABI between the reference's C++ host (chain search,
 * neural generation hooks) and hand-written HIP kernels. It plugs in as one
 * new raster backend value of the versioned neural options
 * (reference include/blitz/blitz.h:135, raster_backend) and consumes the
 * reference's ABI-5 mesh descriptor unchanged (include/blitz/blitz.h:18-37).
 * No C++ exception crosses this boundary; errors map to blitz_status exactly
 * as the reference's C shim does (src/c_api.cpp:187-202) and the message is
 * copied, truncated, into the caller's buffer (src/c_api.cpp:26-32).
 *
 * Threading: one audit object per calling thread; each object owns its HIP
 * stream and workspace. Every entry point restores the caller's current HIP
 * device (reference src/neural/cuda.cuh:32-45).
 */
#ifndef BLITZ_HIP_H
#define BLITZ_HIP_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* The reference's ABI-5 descriptors. When blitz/blitz.h is included first its
 * definitions are used; the layouts below are identical to it. */
#ifndef BLITZ_REMESHER_H
#define BLITZ_ABI_VERSION 5u
typedef enum blitz_status {
    BLITZ_OK = 0,
    BLITZ_INVALID_ARGUMENT = 1,
    BLITZ_OUT_OF_MEMORY = 2,
    BLITZ_INTERNAL_ERROR = 3,
    BLITZ_CANCELLED = 4,
    BLITZ_BUDGET_LIMITED = 5,
    BLITZ_UNAVAILABLE = 6
} blitz_status;
typedef struct blitz_stream {
    const void* data;
    size_t count, stride;
} blitz_stream;
typedef struct blitz_mesh {
    uint32_t struct_size, abi_version;
    blitz_stream positions, normals, uv, colors, tangents;
    const uint32_t* indices;
    size_t index_count;
    const uint16_t* materials;
    size_t material_count;
    const uint8_t* double_sided;
    size_t double_sided_count;
} blitz_mesh;
typedef struct blitz_curve_point {
    float x, y;
} blitz_curve_point;
/* ABI-5 generation settings (reference include/blitz/blitz.h:41-63). */
typedef struct blitz_settings {
    uint32_t struct_size, abi_version;
    uint8_t levels, profile;
    uint16_t triangle_overhead_bps;
    uint32_t max_added_vertex_bytes_bps;
    uint8_t objective, beam_width, search_supersample, audit_supersample;
    uint8_t max_supersample, prune, force_scalar, coupled_wedges;
    uint16_t candidate_budget, search_ortho, search_perspective, audit_ortho, audit_perspective;
    uint32_t search_seed, audit_seed;
    double pixels_per_meter, meters_per_unit, base_pixels, last_pixels, max_lod0_delta_px;
    double normal_weight, color_weight, material_weight;
    double max_changed_area;
    const blitz_curve_point* transition;
    size_t transition_count;
    const blitz_curve_point* normal_importance;
    size_t normal_importance_count;
    const blitz_curve_point* attribute_importance;
    size_t attribute_importance_count;
    int (*cancelled)(void*);
    void* user_data;
} blitz_settings;
#endif

/* New enumerator for blitz_neural_options_v2.raster_backend (CUDA = 0,
 * Vulkan = 1 in the reference; include/blitz/neural.hpp:11). */
#define BLITZ_NEURAL_RASTER_HIP 2u
#define BLITZ_HIP_API_VERSION 1u

/* ---- shared value types --------------------------------------------------- */

/* blitz::Bounds (include/blitz/mesh.hpp:117). */
typedef struct blitz_hip_bounds {
    float center[3];
    double radius;
} blitz_hip_bounds;

/* blitz::Camera (include/blitz/evaluate.hpp:46); cameras() output. */
typedef struct blitz_hip_camera {
    float right[3], up[3], forward[3];
    double distance, focal, scale;
    uint8_t perspective;
} blitz_hip_camera;

/* blitz::EvalSettings (include/blitz/evaluate.hpp:24-33) as a versioned POD.
 * profile: 0 coverage, 1 normals, 2 attributes; every audit entry point
 * implements all three except the bound-reference patch audits, which take
 * coverage only (others return BLITZ_UNAVAILABLE). cancelled is polled once
 * per view in view order, as the reference does (src/evaluate.cpp:319). */
typedef struct blitz_hip_eval_settings {
    uint32_t struct_size, version;
    uint8_t profile, supersample, max_supersample, force_two_sided;
    uint8_t conservative_screen, reserved0[3];
    uint16_t views_orthographic, views_perspective;
    uint32_t views_rotation_seed;
    double screen_size, limit, max_changed_area;
    double normal_weight, color_weight, material_weight;
    int (*cancelled)(void*);
    void* user_data;
} blitz_hip_eval_settings;

/* blitz::Measurement (include/blitz/evaluate.hpp:34-39). */
typedef struct blitz_hip_measurement {
    double error, coverage, coverage_upper, changed_area, normal_degrees;
    uint32_t worst_view, changed_area_worst_view, views_evaluated;
    uint8_t supersample, complete, passed, resource_limited, cancelled;
    uint8_t reserved[3];
} blitz_hip_measurement;

/* Options of one audit object; the HIP counterpart of the subset of
 * blitz::NeuralOptions (include/blitz/neural.hpp:55-86) the audit reads. */
typedef struct blitz_hip_audit_options {
    uint32_t struct_size, version;
    int32_t device;
    uint32_t memory_mib;    /* workspace cap; 0 = no cap */
    uint8_t cache_rasters;  /* reuse the reference's masks across audits of one uploaded reference
                             * (RasterMemo, src/neural/evaluate.cu:386-511); 0 by default */
    uint8_t face_split;     /* reference/candidate face split of coverage rasters over a shared vertex
                             * set: 0 automatic (large rasters), 1 never, 2 whenever possible; the
                             * masks are the same either way */
    uint8_t autotune;       /* 1 (default): time the raster launch shapes on the first launch of each
                             * shape class and keep the fastest; 0: fixed heuristic shapes */
    uint8_t reserved[5];
} blitz_hip_audit_options;

typedef struct blitz_hip_audit blitz_hip_audit;             /* AuditCuda (src/neural/internal.hpp:131) */
typedef struct blitz_hip_device_mesh blitz_hip_device_mesh; /* DeviceMeshView (src/neural/device_view.hpp:16) */

/* Per-call kernel timing of the last evaluate/certify call, for profiling. */
typedef struct blitz_hip_audit_timing {
    uint32_t struct_size;
    uint32_t raster_tasks, distance_views, passes;
    double raster_ms, compare_ms, distance_ms, total_ms;
    uint64_t raster_triangles; /* triangle-view setups issued */
    uint64_t mask_bytes;       /* coverage mask bytes written */
    double project_ms;         /* projection launches alone (device time) */
    double band_ms;            /* band-raster launches alone (device time) */
    uint32_t band_launches;
    uint32_t patch_fallbacks;  /* patches audited by the full path */
    double patch_ms;           /* patch kernel launches (device time) */
    double appearance_ms;      /* attribute rasters, normal angle, attributed distance (device time) */
    uint32_t attributed_views; /* views whose attributed distance was measured */
    uint32_t split_launches;   /* raster launches drawn from a reference/candidate face split */
    double split_ms;           /* face split kernels (device time) */
    uint32_t raster_chunks;    /* face chunks per task of the last raster launch */
    uint32_t tuned_launches;   /* raster launches whose shape was measured in this call */
} blitz_hip_audit_timing;

/* ---- audit ---------------------------------------------------------------- */

uint32_t blitz_hip_api_version(void);
/* Nonzero when a HIP device with gfx950 kernels is usable. */
int blitz_hip_available(int32_t device);
/* Waits for all work on the device (hipDeviceSynchronize); restores the
 * caller's current device. */
blitz_status blitz_hip_device_synchronize(int32_t device);

blitz_status blitz_hip_audit_options_init(blitz_hip_audit_options*, size_t);
/* Defaults of blitz::EvalSettings (include/blitz/evaluate.hpp:24-33). */
blitz_status blitz_hip_eval_settings_init(blitz_hip_eval_settings*, size_t);

/* blitz::bounds (src/mesh.cpp:77-86). */
blitz_status blitz_hip_bounds_of(const blitz_mesh*, blitz_hip_bounds*, char* error, size_t capacity);
/* blitz::cameras (src/evaluate.cpp:108-124), host libm. */
blitz_status blitz_hip_cameras(const blitz_hip_bounds*, double screen, uint16_t orthographic,
                               uint16_t perspective, uint32_t rotation_seed, blitz_hip_camera* out,
                               size_t capacity);

/* AuditCuda(const NeuralOptions&) (src/neural/internal.hpp:139). */
blitz_status blitz_hip_audit_create(const blitz_hip_audit_options*, blitz_hip_audit**, char* error,
                                    size_t capacity);
void blitz_hip_audit_destroy(blitz_hip_audit*);

/* Uploads a validated mesh (blitz::validate, src/mesh.cpp:92-106) into HBM.
 * The handle stays valid until released or its audit is destroyed. */
blitz_status blitz_hip_mesh_upload(blitz_hip_audit*, const blitz_mesh*, blitz_hip_device_mesh**,
                                   char* error, size_t capacity);
void blitz_hip_mesh_release(blitz_hip_device_mesh*);

/* blitz::evaluate_gpu / AuditCuda::evaluate(MeshView, MeshView, ...)
 * (include/blitz/neural.hpp:148-149, src/neural/evaluate.cu:2261-2341):
 * coverage audit with the reference's refinement and first-failing-view
 * semantics (src/evaluate.cpp:312-352). Host meshes are uploaded per call. */
blitz_status blitz_hip_audit_evaluate(blitz_hip_audit*, const blitz_mesh* reference,
                                      const blitz_mesh* candidate, const blitz_hip_bounds*,
                                      const blitz_hip_eval_settings*, blitz_hip_measurement* out,
                                      char* error, size_t capacity);
/* Same, on meshes already resident in HBM (AuditCuda::evaluate(MeshView,
 * DeviceMeshView, ...), src/neural/internal.hpp:163-166). */
blitz_status blitz_hip_audit_evaluate_device(blitz_hip_audit*, const blitz_hip_device_mesh* reference,
                                             const blitz_hip_device_mesh* candidate,
                                             const blitz_hip_bounds*, const blitz_hip_eval_settings*,
                                             blitz_hip_measurement* out, char* error,
                                             size_t capacity);
/* Batched audit of n candidates against one reference, all views of all
 * candidates in shared launches (AuditCuda::certify_candidates,
 * src/neural/evaluate.cu:2413-2578, without its 8-candidate limit). out[i]
 * equals blitz_hip_audit_evaluate_device(reference, candidates[i]). */
blitz_status blitz_hip_audit_evaluate_batch(blitz_hip_audit*, const blitz_hip_device_mesh* reference,
                                            const blitz_hip_device_mesh* const* candidates,
                                            size_t count, const blitz_hip_bounds*,
                                            const blitz_hip_eval_settings*,
                                            blitz_hip_measurement* out, char* error,
                                            size_t capacity);

/* ---- bound reference and local patches ------------------------------------ */

/* A reference bound for patch audits: per view its projection, coverage mask,
 * covered sample count and the reference faces binned per 8x8 tile, kept in
 * HBM (AuditCuda::bind_reference and the reference raster memo,
 * src/neural/internal.hpp:141-144, src/neural/evaluate.cu:386-511). The
 * reference mesh must not carry per-face materials. */
typedef struct blitz_hip_binding blitz_hip_binding;
blitz_status blitz_hip_audit_bind(blitz_hip_audit*, const blitz_mesh* reference, const blitz_hip_bounds*,
                                  const blitz_hip_eval_settings*, blitz_hip_binding**, char* error,
                                  size_t capacity);
void blitz_hip_binding_release(blitz_hip_binding*);

/* Audits `count` local patches of the bound reference in shared launches
 * (certify_candidates with dirty-tile re-rasterization,
 * src/neural/evaluate.cu:216-273,2413-2578). Patch i removes the reference
 * faces removed[removed_offsets[i] .. removed_offsets[i+1]) (ascending ids)
 * and appends the faces added[3*added_offsets[i] .. 3*added_offsets[i+1])
 * (vertex ids of the reference); out[i] equals blitz_hip_audit_evaluate of
 * the reference against that candidate (reference faces minus the region, in
 * order, then the added faces). settings must have the bind's screen size,
 * supersample, view set and sidedness; limit, max_changed_area and
 * max_supersample may differ. *fallbacks (optional) receives the number of
 * patches audited by the full path (clipped views, refinement, regions over
 * 512 faces or 1024 dirty tiles). */
blitz_status blitz_hip_audit_evaluate_patches(blitz_hip_audit*, blitz_hip_binding*, size_t count,
                                              const uint32_t* removed_offsets, const uint32_t* removed,
                                              const uint32_t* added_offsets, const uint32_t* added,
                                              const blitz_hip_eval_settings*, blitz_hip_measurement* out,
                                              uint32_t* fallbacks, char* error, size_t capacity);

/* Seeded local patches of a mesh (host): `count` breadth-first face regions
 * of uniform size in [min_faces, max_faces], each re-meshed by legal interior
 * edge collapses onto existing vertices. Outputs CSR arrays as taken by
 * blitz_hip_audit_evaluate_patches; *removed_count and *added_count (faces)
 * receive the sizes needed, and null output arrays only query them. */
blitz_status blitz_hip_local_patches(const blitz_mesh*, uint32_t count, uint64_t seed, uint32_t min_faces,
                                     uint32_t max_faces, uint32_t* removed_offsets, uint32_t* removed,
                                     size_t removed_capacity, uint32_t* added_offsets, uint32_t* added,
                                     size_t added_capacity, size_t* removed_count, size_t* added_count,
                                     char* error, size_t capacity);

/* ---- collapse actions ------------------------------------------------------ */

/* Device state of one mesh's collapse actions: the GPU action topology of
 * src/neural/action_gpu.cu:178-343 with the semantics of the CPU
 * blitz::neural::ActionState (src/neural/action.cpp:27-145). Vertices with
 * bit-equal positions weld to the smallest index; an optional uv stream in the
 * mesh descriptor enables the UV orientation test. */
typedef struct blitz_hip_actions blitz_hip_actions;
blitz_status blitz_hip_actions_create(int32_t device, const blitz_mesh*, blitz_hip_actions**, char* error,
                                      size_t capacity);
void blitz_hip_actions_destroy(blitz_hip_actions*);
/* Undirected edges, legal directed actions and the device build time. */
blitz_status blitz_hip_actions_info(const blitz_hip_actions*, uint32_t* edges, uint32_t* legal, double* build_ms);
/* Legal directed collapses (from, to pairs) in ActionState::actions order
 * (src/neural/action.cpp:371-383: edges ascending, a->b then b->a), the
 * enumerate/legal kernels of src/neural/action_gpu.cu:487-643. */
blitz_status blitz_hip_actions_legal(blitz_hip_actions*, uint32_t* pairs, size_t capacity, char* error,
                                     size_t error_capacity);
/* ActionState::independent (src/neural/action.cpp:280-304) / select_batch
 * (src/neural/action_gpu.cu:865-899): greedy non-overlapping footprints over
 * `order` (indices of legal actions), at most `maximum` (1..64) actions and
 * `maximum_removed` faces; writes the selected indices. */
blitz_status blitz_hip_actions_independent(blitz_hip_actions*, const uint32_t* order, size_t n, uint32_t maximum,
                                           uint64_t maximum_removed, uint32_t* selected, size_t* selected_count,
                                           char* error, size_t capacity);
/* ActionState::trial (src/neural/action.cpp:308-344) / keep_faces + copy_trial
 * (src/neural/action_gpu.cu:757-792): the index buffer with the legal actions
 * `rows` (1..64, non-overlapping) applied. */
blitz_status blitz_hip_actions_trial(blitz_hip_actions*, const uint32_t* rows, size_t n, uint32_t* indices,
                                     size_t index_capacity, size_t* index_count, char* error, size_t capacity);
/* Each legal action rows[i] as a patch for blitz_hip_audit_evaluate_patches:
 * the faces around its `from` vertex (ascending) removed and their remapped
 * survivors added. Null output arrays only query the sizes (faces). */
blitz_status blitz_hip_actions_patches(blitz_hip_actions*, const uint32_t* rows, size_t n,
                                       uint32_t* removed_offsets, uint32_t* removed, size_t removed_capacity,
                                       uint32_t* added_offsets, uint32_t* added, size_t added_capacity,
                                       size_t* removed_count, size_t* added_count, char* error, size_t capacity);

/* blitz::neural::graph (src/neural/graph.cpp:35-187) of the state's mesh:
 * the 24 vertex channels (position over the bounds diameter, unit normal,
 * colour, UV, tangent, stream presence, seam / boundary / material flags,
 * log degree, log mean edge) per vertex, vertex-major; `capacity` in floats
 * (at least 24 per vertex). The normals, colours and tangents streams are
 * the ones given to blitz_hip_actions_create. */
blitz_status blitz_hip_actions_vertex_features(blitz_hip_actions*, float* out, size_t capacity, char* error,
                                               size_t error_capacity);
/* ActionState::actions (src/neural/action.cpp:371-450; fill_action_features,
 * src/neural/action_gpu.cu:644-730): the feature row of every legal action,
 * in legal order, `width` floats each: 80 (the action scorer's input), or
 * 128 for placement rows (channels 80..127: the second wedge's source and
 * target vertex channels for two-wedge actions). `conditions` are the eight
 * neural::condition values (src/neural/graph.cpp:253-263), which must be
 * finite. `input_triangles` is the triangle count of the input the state
 * descends from (channel 66 = current / input triangles); 0 takes the
 * state's own. `capacity` in floats. */
blitz_status blitz_hip_actions_features(blitz_hip_actions*, const float* conditions, uint32_t width,
                                        uint64_t input_triangles, float* rows, size_t capacity, char* error,
                                        size_t error_capacity);

/* Action and placement policies: ActionCuda (src/neural/action.hpp:70-87,
 * src/neural/inference.cu:183-290), three compensated FP32 layers (ReLU after
 * the first two; the reference's compensated_linear_kernel arithmetic, so the
 * outputs are bit-identical to it). architecture 2: 80 inputs, 3 outputs
 * (rank, source, adjacent logits); 3 and 4: 128 inputs, 12 outputs
 * (src/neural/placement.hpp:8-31). hidden_width 64, 128 or 256; weights
 * float32 [out,in] then bias per layer, count = (in+1)*h + (h+1)*h +
 * (h+1)*out, all finite. */
typedef struct blitz_hip_policy blitz_hip_policy;
blitz_status blitz_hip_policy_create(int32_t device, uint32_t architecture, uint32_t hidden_width,
                                     const float* weights, size_t count, blitz_hip_policy**, char* error,
                                     size_t capacity);
void blitz_hip_policy_destroy(blitz_hip_policy*);
/* ActionCuda::predict: `rows` input rows (host, finite) to rows x outputs
 * floats; a non-finite output is BLITZ_INTERNAL_ERROR. */
blitz_status blitz_hip_policy_predict(blitz_hip_policy*, const float* input, size_t rows, float* output,
                                      size_t output_capacity, char* error, size_t capacity);
/* Learned ranking of the reference's GPU action state (rank with a policy,
 * src/neural/action_gpu.cu:1751-1798 and rank_actions 809-844): every legal
 * action's feature row (blitz_hip_actions_features at the policy's input
 * width; architecture 4 stores preserve_uv in feature 79) through the
 * policy, ordered by its first output, descending, ties in enumeration
 * order; the order stays on the device for later batch selection and is
 * written to `order` (legal-action indices) when it is not null. One policy
 * serves one thread at a time. */
blitz_status blitz_hip_actions_rank_policy(blitz_hip_actions*, blitz_hip_policy*, const float* conditions,
                                           uint64_t input_triangles, int32_t preserve_uv, uint32_t* order,
                                           size_t capacity, char* error, size_t error_capacity);

/* Placement decoding (decode_placements / default_placements,
 * src/neural/action_gpu.cu:1157-1200; grid_position 380-390): the
 * trajectory's quantization box (positions inside it snap to its unorm16
 * grid, src/neural/vertex_storage.hpp:24-32) and the bitmap of exact-position
 * vertices (an edit touching one never snaps). */
typedef struct blitz_hip_placement_options {
    uint32_t struct_size;
    int32_t preserve_uv;          /* architecture 4's feature 79 */
    int32_t quantized;            /* snap to the grid of [low, low + extent] */
    float low[3], extent[3];
    const uint32_t* exact_bits;   /* bit i of word i / 32: vertex i is exact; or null */
    size_t exact_words;
} blitz_hip_placement_options;
blitz_status blitz_hip_placement_options_init(blitz_hip_placement_options*, size_t size);
/* 9 floats per legal action: the placed position, then the unit normal of
 * each destination wedge (zeros without normals or a second wedge). With a
 * placement policy (architecture 3 or 4): the edge midpoint plus the policy's
 * clamped offset (outputs 3..5, in edge lengths), snapped; normals plus the
 * predicted residuals (outputs 6..11) when any of conditions[3..5] is
 * non-zero, as the reference supervises them. Without a policy
 * (policy = NULL): the destination vertex and its wedges' unit normals. */
blitz_status blitz_hip_actions_placements(blitz_hip_actions*, blitz_hip_policy*, const float* conditions,
                                          uint64_t input_triangles, const blitz_hip_placement_options*,
                                          float* out, size_t capacity, char* error, size_t error_capacity);
/* Wedge mapping of every legal action (the remap of ActionState::mapping,
 * src/neural/action.cpp:146-215): 5 words each, the wedge count then the
 * (from, to) vertex index pairs in ascending order, 0xFFFFFFFF for an absent
 * second pair. */
blitz_status blitz_hip_actions_mapping(blitz_hip_actions*, uint32_t* out, size_t capacity, char* error,
                                       size_t error_capacity);

/* Neural model files: the reference's BLZNET01/02/03 format (load_weights /
 * save_weights, src/neural/model_io.cpp:61-153): header, provenance text,
 * float32 weights and a SHA-256 checksum trailer (64 hex digits). Loading
 * checks size, magic, architecture / width / weight count, use scope,
 * payload length and checksum, and rejects non-finite weights
 * (BLITZ_INVALID_ARGUMENT, as the reference). */
typedef struct blitz_hip_model blitz_hip_model;
blitz_status blitz_hip_model_load(const char* path, blitz_hip_model**, char* error, size_t capacity);
void blitz_hip_model_destroy(blitz_hip_model*);
/* blitz_neural_model_sha256 (include/blitz/blitz.h:163-164): SHA-256 of the
 * whole file, 64 lower-case hex digits, borrowed until destruction. */
const char* blitz_hip_model_sha256(const blitz_hip_model*);
typedef struct blitz_hip_model_info {
    uint32_t struct_size;
    uint32_t architecture;    /* 1 graph encoder, 2 action policy, 3 / 4 placement policies */
    uint32_t hidden_width;    /* 64, 128 or 256 */
    uint32_t use;             /* 0 unrestricted, 1 endpoint-reuse only (architecture 4) */
    uint64_t weight_count;
    const float* weights;     /* borrowed */
    uint64_t provenance_size;
    const char* provenance;   /* borrowed, not NUL-terminated */
} blitz_hip_model_info;
blitz_status blitz_hip_model_info_of(const blitz_hip_model*, blitz_hip_model_info*);
/* save_weights: BLZNET01 unless the width is not 64 (BLZNET02) or the use
 * is restricted (BLZNET03); written to a temporary file renamed over `path`. */
blitz_status blitz_hip_model_save(const char* path, uint32_t architecture, uint32_t hidden_width, uint32_t use,
                                  const float* weights, size_t count, const char* provenance,
                                  size_t provenance_size, char* error, size_t capacity);
/* The policy of a loaded action (2) or placement (3, 4) model. */
blitz_status blitz_hip_policy_from_model(int32_t device, const blitz_hip_model*, blitz_hip_policy**, char* error,
                                         size_t capacity);

/* Work counters of blitz::neural::ActionStats (src/neural/action.hpp:62-67). */
typedef struct blitz_hip_action_stats {
    uint32_t struct_size;
    uint32_t states; /* action states built on the device */
    uint64_t ranked, trials, accepted, rejected;
    double ms;       /* wall time of the whole call */
    /* wall time of the stages: state builds, ranking, batch selection,
     * trials, gate audits */
    double build_ms, rank_ms, select_ms, trial_ms, gate_ms;
} blitz_hip_action_stats;
/* blitz::neural::execute_actions (src/neural/action.cpp:453-525; the GPU
 * executor of src/neural/action_gpu.cu, checked against it by
 * tests/neural_action_gpu.cpp:140-190): collapse `input` towards `target`
 * triangles within `budget` trials, ranking by `ranking` (0 learned: the BF16
 * logits of `scorer` on each mesh state's graph; 1 constant, 3 shortest edge,
 * 4 current plane; `scorer` may be null except for 0), batches of up to
 * `batch` (1..64)
 * independent actions, each trial gated by an exact audit against `input`
 * under `gate` (accepted when the Measurement passes). Writes the final index
 * buffer. Needs the blitz_hip_audit of the device the actions run on. */
typedef struct blitz_hip_scorer blitz_hip_scorer;
blitz_status blitz_hip_execute_actions(blitz_hip_audit*, const blitz_mesh* input, const blitz_hip_bounds*,
                                       const blitz_hip_eval_settings* gate, uint64_t target, uint32_t budget,
                                       uint8_t ranking, blitz_hip_scorer* scorer, uint8_t batch, uint32_t* indices,
                                       size_t index_capacity, size_t* index_count, blitz_hip_action_stats* stats,
                                       char* error, size_t capacity);

/* ---- LOD chain ------------------------------------------------------------- */

/* The fields of blitz_neural_options_v2 (include/blitz/blitz.h:127-146) that
 * the HIP chain reads, plus the explicit output mode. Supported: origin
 * BLITZ_NEURAL_SOURCE (0), output reuse (1: every LOD indexes the source's
 * vertices; research.output = Reuse), GPU confirmation (1). Other values
 * return BLITZ_UNAVAILABLE. */
typedef struct blitz_hip_chain_options {
    uint32_t struct_size, version;
    uint32_t action_trials;     /* 64: trials per proposal */
    uint8_t action_batch;       /* 32: independent collapses per trial, 1..64 */
    uint8_t ranking;            /* 4: 0 learned (scorer), 1 constant, 3 shortest edge, 4 current plane */
    uint8_t origin;             /* 0: source */
    uint8_t output;             /* 1: reuse */
    uint8_t confirmation;       /* 1: GPU */
    uint8_t overdraw_tiebreak;  /* 1: equal-cost finalists ordered by the top-left overlap */
    uint8_t reserved[6];
} blitz_hip_chain_options;

/* One scheduled level (blitz_lod_info, include/blitz/blitz.h:64-72): the
 * triangles as indices into the source's vertices (borrowed from the chain,
 * valid until it is destroyed), the schedule and both audits. */
typedef struct blitz_hip_chain_lod {
    uint32_t struct_size;
    const uint32_t* indices;
    size_t index_count;
    const uint16_t* materials;
    size_t material_count;
    double screen_pixels, transition_limit, source_limit;
    blitz_hip_measurement adjacent, source;
} blitz_hip_chain_lod;

/* Counters of the generation (blitz::Result, include/blitz/remesher.hpp:119-151,
 * and blitz_neural_info, include/blitz/blitz.h:153-158). legal_collapses
 * counts collapses that passed all four gates (source search, adjacent
 * search, source audit, adjacent audit). */
typedef struct blitz_hip_chain_info {
    uint32_t struct_size;
    uint32_t levels, status, fallback_levels;
    uint64_t candidate_evaluations;
    uint64_t audit_evaluations[4], rejected_gates[4], area_rejected_gates[4];
    uint64_t decoded, legal_collapses, rejected_collapses, action_ranked, action_trials, reference_rejections;
    uint32_t bounded_audits, finalists, selected, proposals;
    uint64_t gate_audits, confirm_audits;
    double ms, executor_ms, gate_ms, offer_ms, confirm_ms, tiebreak_ms;
    /* executor stages over all proposals: state builds, rankings, batch
     * selection, trials, and next-state builds run while the gates audit */
    double build_ms, rank_ms, select_ms, trial_ms, speculative_ms;
    double speculative_build_ms; /* of speculative_ms: the state builds */
} blitz_hip_chain_info;

/* One proposal (blitz_neural_action_proposal, include/blitz/blitz.h:147-152)
 * and the gate its offer stopped at (0 accepted, 1..4 the four gates,
 * 5 invalid, 6 growth). */
typedef struct blitz_hip_chain_proposal {
    uint32_t struct_size;
    uint32_t level, parent_triangles, start_triangles, target_triangles, final_triangles, trials,
        accepted_batches;
    uint64_t accepted, rejected;
    double screen_pixels;
    uint8_t stop_reason, gate, reserved[6];
} blitz_hip_chain_proposal;

typedef struct blitz_hip_chain blitz_hip_chain;

blitz_status blitz_hip_chain_options_init(blitz_hip_chain_options*, size_t);
/* blitz_generate_neural (include/blitz/blitz.h:165-167; src/neural/generate.cpp:126-577
 * with detail::generate_with_hooks, src/chain.cpp:1175): the LOD chain of
 * `source` under the ABI-5 settings, every proposal made by the device
 * collapse executor (scorer needed for ranking 0) and audited on the GPU.
 * Returns BLITZ_BUDGET_LIMITED or BLITZ_CANCELLED with a valid chain, as the
 * reference's C shim does (src/c_api.cpp:180-186). */
blitz_status blitz_hip_generate(blitz_hip_audit*, const blitz_mesh* source, const blitz_settings*,
                                const blitz_hip_chain_options*, blitz_hip_scorer* scorer, blitz_hip_chain** out,
                                char* error, size_t capacity);
void blitz_hip_chain_destroy(blitz_hip_chain*);
/* blitz::schedule (src/chain.cpp:101-118), host only: per level the screen
 * size, transition limit and source limit (3 doubles each, `levels` rows). */
blitz_status blitz_hip_schedule(const blitz_settings*, const blitz_hip_bounds*, double* out, size_t capacity,
                                char* error, size_t error_capacity);
blitz_status blitz_hip_chain_info_of(const blitz_hip_chain*, blitz_hip_chain_info*);
blitz_status blitz_hip_chain_lod_of(const blitz_hip_chain*, size_t level, blitz_hip_chain_lod*);
/* Triangles per level of finalist i (levels entries). */
blitz_status blitz_hip_chain_finalist(const blitz_hip_chain*, size_t i, uint32_t* triangles, size_t capacity);
blitz_status blitz_hip_chain_proposal_of(const blitz_hip_chain*, size_t i, blitz_hip_chain_proposal*);

/* blitz::detail::rasterize_coverage (src/coverage.hpp:12, src/evaluate.cpp:125-189)
 * for one camera: row-major packed words (index = y*size + x), padding bits
 * zero. *size_out receives the raster side; returns BLITZ_BUDGET_LIMITED when
 * the raster exceeds the 64M-sample cap (src/evaluate.cpp:129). */
blitz_status blitz_hip_raster_coverage(blitz_hip_audit*, const blitz_hip_device_mesh*,
                                       const blitz_hip_bounds*, const blitz_hip_camera*,
                                       double screen, uint8_t supersample, uint8_t force_two_sided,
                                       uint64_t* words, size_t word_capacity, uint32_t* size_out,
                                       uint8_t* clipped, char* error, size_t capacity);
/* The attributed instantiation of blitz::detail::rasterize (src/evaluate.cpp:125-189,
 * Pixel of include/blitz/evaluate.hpp:40-44) for one camera, row-major
 * (index = y*size + x): covered and visible flags, the stored float depth
 * (+inf where nothing is visible), normal xyz, colour rgba and material per
 * sample; each array holds size*size entries (normal 3, colour 4 per sample)
 * and `sample_capacity` is the entry count of the byte arrays. Returns
 * BLITZ_BUDGET_LIMITED over the 64M-sample cap, with *size_out set. */
blitz_status blitz_hip_raster_pixels(blitz_hip_audit*, const blitz_hip_device_mesh*,
                                     const blitz_hip_bounds*, const blitz_hip_camera*,
                                     double screen, uint8_t supersample, uint8_t force_two_sided,
                                     uint8_t* covered, uint8_t* visible, float* depth, float* normal,
                                     float* color, uint16_t* material, size_t sample_capacity,
                                     uint32_t* size_out, uint8_t* clipped, char* error, size_t capacity);
/* blitz::overlap_cuda (include/blitz/neural.hpp:150-151, src/neural/evaluate.cu:2665-2685):
 * the mean over the views {orthographic, perspective, seed} of the overdraw
 * samples per covered pixel, counting every drawable face whose interior holds
 * a pixel centre under the top-left fill rule, at one sample per pixel;
 * +inf when a face of any view lies behind the camera or out of frame. */
blitz_status blitz_hip_overlap(blitz_hip_audit*, const blitz_hip_device_mesh*, const blitz_hip_bounds*,
                               double screen, uint16_t orthographic, uint16_t perspective, uint32_t seed,
                               double* out, char* error, size_t capacity);
/* blitz::detail::coverage_distance (src/coverage.hpp:13, src/evaluate.cpp:196-216)
 * on two row-major packed masks of side `size`. */
blitz_status blitz_hip_coverage_distance(blitz_hip_audit*, const uint64_t* a, const uint64_t* b,
                                         uint32_t size, uint8_t supersample, double* out,
                                         char* error, size_t capacity);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): blitz_hip_coverage_distance on masks already in device
 * memory, with every temporary in one fixed array inside the library (16 MiB
 * per GPU, held until the process ends). `a` and `b` are device pointers to
 * row-major packed masks (bit y*size + x, (size*size + 63) / 64 words each);
 * the distance is written to the device double `out`. Work is queued on
 * `stream` (a hipStream_t; null for the default stream) and the call returns
 * without waiting. size <= 8192. Calls that overlap on one GPU share the
 * array and race; calls on one stream are ordered. */
blitz_status blitz_hip_coverage_distance_noalloc(int32_t device, const uint64_t* a, const uint64_t* b,
                                                 uint32_t size, uint8_t supersample, double* out, void* stream,
                                                 char* error, size_t capacity);

blitz_status blitz_hip_audit_last_timing(const blitz_hip_audit*, blitz_hip_audit_timing*);

/* No counterpart in the reference (which calls CUB): the library's own device
 * sort, scan and run-length primitives (device_sort.hip) on host arrays, for
 * tests of the primitives themselves. kind 0: stable ascending sort of
 * uint32 keys by bits [begin_bit, end_bit), values carried (input positions
 * when `values` is null); 1: the same for uint64 keys; 2: stable descending
 * sort of float keys (-0 equal to +0); 3: exclusive sum of uint32; 4:
 * inclusive max of uint32; 5: run-length encoding of uint64 keys (unique keys
 * to keys_out, counts to values_out, run count to *runs_out); 6: exclusive
 * sum of uint64. Outputs hold n elements. */
blitz_status blitz_hip_device_primitive(int32_t device, uint32_t kind, const void* keys, const uint32_t* values,
                                        uint32_t n, int32_t begin_bit, int32_t end_bit, void* keys_out,
                                        uint32_t* values_out, uint32_t* runs_out, char* error, size_t capacity);

/* ---- classical proposer --------------------------------------------------- */

/* Quadric edge collapse with reused vertices: the classical proposal
 * (blitz::reduce with OutputMode::Reuse, src/simplify.cpp:337-367; greedy
 * heap order). Writes at most `target_triangles` triangles (fewer collapses
 * if no legal one remains) as indices into the input vertex array. Host CPU. */
blitz_status blitz_hip_reduce_reuse(const blitz_mesh*, size_t target_triangles, uint32_t* indices_out,
                                    size_t index_capacity, size_t* index_count, char* error,
                                    size_t capacity);

/* blitz::reduce with OutputMode::Reuse and the default quadric objective
 * (src/simplify.cpp:43-329, reduce :337-367) on the device: the pass
 * structure, candidate costs, (cost, u, v) order, link and fold tests and
 * budget of the reference, every pass's walk as a greedy independent set in
 * rank order. Same triangles, in the same order, as the reference. Reads
 * positions, indices, materials and normals. `report` (may be null) receives
 * passes, greedy rounds and collapses. */
typedef struct blitz_hip_reduce_report {
    uint32_t struct_size;
    uint32_t passes, rounds, tail_collapses;
    uint64_t collapses;
} blitz_hip_reduce_report;
blitz_status blitz_hip_reduce_reuse_device(int32_t device, const blitz_mesh*, size_t target_triangles,
                                           uint32_t* indices_out, size_t index_capacity, size_t* index_count,
                                           blitz_hip_reduce_report* report, char* error, size_t capacity);

/* ---- edge-graph proposal scorer ------------------------------------------ */

/* The HIP scorer ranks every directed contraction (u -> v) of a mesh: 32
 * vertex features, three mean-aggregation message-passing layers of width 128
 * and an edge MLP [h_u ; h_v] 256 -> 128 -> 4 (logit, placement xyz). It takes
 * the place of ActionCuda::predict_device and the ranking step of
 * GpuActionState (src/neural/action.hpp:75, src/neural/action_gpu.cu:1751-1798).
 * Weights: float32, row-major [out, in] then bias, layer after layer
 * (src/neural/internal.hpp:42-43). */
typedef struct blitz_hip_scorer blitz_hip_scorer;
typedef struct blitz_hip_graph blitz_hip_graph;

/* Rank order of the legal actions (rank_actions + SortPairsDescending,
 * src/neural/action_gpu.cu:809-844,1793-1795): `ranking` takes
 * blitz_neural_options_v2.ranking values, 0 learned (the logits of a scorer
 * graph of the same mesh after blitz_hip_scorer_forward), 1 constant,
 * 3 shortest edge, 4 current plane; 2 (shuffled) returns BLITZ_UNAVAILABLE.
 * Ties keep ascending action order. */
blitz_status blitz_hip_actions_rank(blitz_hip_actions*, uint8_t ranking, blitz_hip_scorer* scorer,
                                    blitz_hip_graph* graph, uint32_t* order, size_t capacity, char* error,
                                    size_t error_capacity);

#define BLITZ_HIP_PRECISION_FP32 0u /* deterministic compensated FP32 */
#define BLITZ_HIP_PRECISION_BF16 1u /* BF16 MFMA, FP32 accumulation */

size_t blitz_hip_scorer_weight_count(void);
blitz_status blitz_hip_scorer_create(int32_t device, const float* weights, size_t count, blitz_hip_scorer**,
                                     char* error, size_t capacity);
void blitz_hip_scorer_destroy(blitz_hip_scorer*);
blitz_status blitz_hip_scorer_set_weights(blitz_hip_scorer*, const float* weights, size_t count, char* error,
                                          size_t capacity);
/* Builds the mesh graph (neighbour CSR, one action per directed edge in CSR
 * order) and its vertex features on the device. */
blitz_status blitz_hip_graph_create(blitz_hip_scorer*, const blitz_mesh*, const blitz_hip_bounds*,
                                    blitz_hip_graph**, char* error, size_t capacity);
void blitz_hip_graph_destroy(blitz_hip_graph*);
/* Disjoint union of graphs of one scorer, in the order given: vertex,
 * action and face ids of each part follow those of the parts before it, and
 * each part keeps the vertex features of the bounds it was built with, so
 * one forward pass or training batch spans several meshes. The reference
 * trains on many meshes through its resident dataset
 * (training/resident.hpp:344-488); this is the device side of that for the
 * edge-graph scorer. The parts stay valid and independent. */
blitz_status blitz_hip_graph_concat(blitz_hip_scorer*, blitz_hip_graph* const* parts, size_t count,
                                    blitz_hip_graph** out, char* error, size_t capacity);
blitz_status blitz_hip_graph_info(const blitz_hip_graph*, uint32_t* vertices, uint32_t* actions);
blitz_status blitz_hip_graph_features(blitz_hip_scorer*, const blitz_hip_graph*, float* out, size_t capacity);
blitz_status blitz_hip_graph_actions(const blitz_hip_graph*, uint32_t* src, uint32_t* dst, size_t capacity);
/* Forward pass over every action; outputs (actions x 4) stay on the device
 * and are copied to `outputs` when it is not null. */
blitz_status blitz_hip_scorer_forward(blitz_hip_scorer*, blitz_hip_graph*, uint8_t precision, float* outputs,
                                      size_t capacity, char* error, size_t capacity_error);
/* Descending score order of every action (rank_actions + radix sort,
 * src/neural/action_gpu.cu:809-844,1793-1795); `ranking` takes the values of
 * blitz_neural_options_v2.ranking: 0 learned, 1 constant, 3 shortest edge,
 * 4 current plane (the plain classical order); 2 (shuffled) is unavailable.
 * Equal scores keep ascending action order. */
blitz_status blitz_hip_scorer_rank(blitz_hip_scorer*, blitz_hip_graph*, uint8_t ranking, uint32_t* order,
                                   size_t capacity, char* error, size_t error_capacity);
/* Placements of the given actions (decode_placements,
 * src/neural/action_gpu.cu:1157-1188, without grid snapping): float xyz,
 * middle + clamp(y, -1, 1) * edge length in the reference's Vec3 arithmetic.
 * A non-finite prediction returns BLITZ_INTERNAL_ERROR. */
blitz_status blitz_hip_scorer_decode(blitz_hip_scorer*, blitz_hip_graph*, const uint32_t* actions, size_t count,
                                     float* placements, char* error, size_t capacity);
/* Teacher placement candidates of the given actions (teacher_placements,
 * src/neural/action_gpu.cu:1304-1373, positions only, without grid
 * snapping): 10 float xyz per action (from, to, middle, plane-quadric optimum
 * of the faces around both endpoints solved in FP64, middle +/- a quarter
 * edge along x, y, z) and the clamped offset targets of each. */
#define BLITZ_HIP_TEACHER_CANDIDATES 10u
blitz_status blitz_hip_scorer_teacher(blitz_hip_scorer*, blitz_hip_graph*, const uint32_t* actions, size_t count,
                                      float* positions, float* targets, char* error, size_t capacity);
blitz_status blitz_hip_scorer_last_forward_ms(const blitz_hip_scorer*, double* ms);

/* ---- scorer training ----------------------------------------------------- */

/* One optimizer step of the scorer on a batch of labelled actions: forward
 * with every activation kept (FP32 compensated or BF16 MFMA), loss, FP32
 * backward, global-norm clip and AdamW on the FP32 master weights, then the
 * BF16 copy refreshed. Takes the place of the reference's update step
 * (training/update_cuda.hpp:47-53 settings, training/update_cuda.cu:256-393
 * loss, clip and AdamW).
 *
 * Batch: `count` action ids of the graph in pools of `pool` consecutive
 * entries. labels[i]: bit 0 valid, bit 1 preferred (ranking target within
 * its pool), bit 2 placement target present; targets: 3 floats per entry
 * (null when no entry is placed). Loss: pairwise hinge on the logit between
 * preferred and other valid entries of a pool / pairs, logit L2 penalty x
 * penalty / valid / 3, smooth-L1 placement x placement_weight / (3 placed). */
typedef struct blitz_hip_train_settings {
    float margin;           /* 1 */
    float penalty;          /* 1e-4 */
    float lr;               /* 1e-3, in (0, 1] */
    float decay;            /* 1e-4 (decoupled weight decay) */
    float max_norm;         /* 1 (global gradient norm clip) */
    float placement_weight; /* 1 */
    uint32_t pool;          /* 32, 1..256 */
    uint8_t precision;      /* BLITZ_HIP_PRECISION_BF16 */
} blitz_hip_train_settings;

typedef struct blitz_hip_train_result {
    float loss;
    float grad_norm;  /* before clipping */
    uint8_t applied;  /* 0 when loss or gradient norm is not finite (no update) */
    uint32_t step;    /* optimizer steps applied so far */
} blitz_hip_train_result;

void blitz_hip_train_settings_init(blitz_hip_train_settings*);
blitz_status blitz_hip_scorer_train_step(blitz_hip_scorer*, blitz_hip_graph*, const uint32_t* actions,
                                         const uint8_t* labels, const float* targets, size_t count,
                                         const blitz_hip_train_settings*, blitz_hip_train_result*, char* error,
                                         size_t capacity);
/* Gradient of the last step (before clipping) and current FP32 weights,
 * blitz_hip_scorer_weight_count() floats each. */
blitz_status blitz_hip_scorer_gradients(blitz_hip_scorer*, float* out, size_t capacity);
blitz_status blitz_hip_scorer_weights(blitz_hip_scorer*, float* out, size_t capacity);
/* Clears the Adam moments and the step counter. */
blitz_status blitz_hip_scorer_reset_optimizer(blitz_hip_scorer*);
blitz_status blitz_hip_scorer_last_step_ms(const blitz_hip_scorer*, double* ms);

#ifdef __cplusplus
}
#endif
#endif
