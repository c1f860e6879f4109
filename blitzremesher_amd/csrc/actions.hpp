// Collapse actions on the device: the counterpart of the reference's GPU
// action state (src/neural/action_gpu.cu:178-343 topology, 487-643 legality,
// 757-808 trial edits, 865-899 batch selection) with the CPU semantics of
// blitz::neural::ActionState (src/neural/action.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <memory>
#include <vector>

#include "audit.hpp"
#include "device_buffer.hpp"
#include "scorer.hpp"

namespace blitz_hip {

class Policy;

// One legal directed half-edge collapse `from` -> `to` (geometric vertices:
// the smallest index among bit-equal positions) and its wedge mapping: in
// every face holding both, index wedge_from[i] becomes wedge_to[i].
struct LegalAction {
    uint32_t from, to;
    uint32_t wedge_from[2], wedge_to[2];
    uint32_t wedges;   // 1 or 2
    uint32_t removed;  // faces the collapse removes (1 boundary, 2 interior)
};

// Device views of one mesh's action topology (all device pointers).
struct ActionTopo {
    const float4* positions;
    const float2* uv;  // or null
    const uint32_t* indices;
    const uint16_t* materials;  // or null
    const uint32_t* geometry;   // geometric vertex of each index
    const uint8_t* invalid;
    const uint8_t* boundary;
    const uint32_t* offsets;       // neighbour lists (sorted)
    const uint32_t* neighbors;
    const uint32_t* face_offsets;  // faces around each geometric vertex (ascending)
    const uint32_t* faces;
    uint32_t vertices, triangles;
    const uint4* face_geo;  // sorted geometric vertices of each face
};

// Host attribute streams of a state's input for the graph features (null
// when absent): normals xyz, colours RGBA8 and tangents xyzw per vertex.
struct VertexStreams {
    const float* normals = nullptr;
    const uint8_t* colors = nullptr;
    const float* tangents = nullptr;
};

// Placement decoding options (src/neural/device_view.hpp:18-33): the
// trajectory's quantization box (positions snap to its unorm16 grid) and the
// host bitmap of exact-position vertices (edits touching one never snap).
struct PlacementOptions {
    bool preserve_uv = false;  // architecture 4's feature 79
    bool quantized = false;
    float low[3] = {0, 0, 0}, extent[3] = {0, 0, 0};
    const uint32_t* exact = nullptr;
    size_t exact_words = 0;
};

class ActionSet {
  public:
    // Builds the topology of `m` (optional per-vertex UVs, 2 floats each) and
    // enumerates its legal actions, edges ascending, a->b then b->a
    // (src/neural/action.cpp:371-383).
    ActionSet(int device, const HostMesh& m, const float* uv);
    ~ActionSet();
    // The same state of another mesh (of the next executor state), keeping
    // the device buffers and, when the positions are unchanged, the weld.
    void rebuild(const HostMesh& m, const float* uv);
    ActionSet(const ActionSet&) = delete;
    ActionSet& operator=(const ActionSet&) = delete;

    uint32_t legal_count() const { return legal_count_; }
    uint32_t edge_count() const { return edges_; }
    void download_legal(uint32_t* pairs) const;  // from, to per legal action
    // ActionState::independent (src/neural/action.cpp:280-304): greedy batch
    // of non-overlapping footprints in `order` (indices of legal actions).
    std::vector<uint32_t> independent(const uint32_t* order, size_t n, uint32_t maximum, uint64_t removable);
    // Rank order of the legal actions (src/neural/action_gpu.cu:809-844,
    // 1793-1795; ranking values of blitz_neural_options_v2): 0 learned (the
    // scorer logits of `graph`), 1 constant, 3 shortest edge, 4 current plane;
    // kept on the device for independent_ranked, copied to order_out if given.
    void rank(uint32_t mode, double diameter, const Graph* graph, uint32_t* order_out);
    // independent() over the device rank order from `position`; *first
    // receives the rank position of the first pick.
    std::vector<uint32_t> independent_ranked(size_t position, uint32_t maximum, uint64_t removable, size_t* first);
    // ActionState::trial (src/neural/action.cpp:308-344): the index buffer of
    // the mesh with the legal actions `rows` applied (and the kept faces'
    // materials, when the mesh has materials).
    std::vector<uint32_t> trial(const uint32_t* rows, size_t n, std::vector<uint16_t>* materials = nullptr);
    // Each legal action `rows[i]` as a patch of the source: the faces around
    // `from` (ascending) removed, their remapped survivors added.
    void patches(const uint32_t* rows, size_t n, std::vector<uint32_t>& removed_offsets,
                 std::vector<uint32_t>& removed, std::vector<uint32_t>& added_offsets,
                 std::vector<uint32_t>& added);
    // blitz::neural::graph (src/neural/graph.cpp:35-187) of the state's
    // mesh: 24 channels per vertex, vertex-major.
    void vertex_features(const VertexStreams& streams, std::vector<float>& out);
    // ActionState::actions (src/neural/action.cpp:371-450): `width` (80, or
    // 128 with the second wedge's vertex channels in 80..127,
    // src/neural/action_gpu.cu:725-729) floats per legal action, in legal
    // order, for the eight conditions; input_faces is the triangle count of
    // the input the state descends from (0: this state's own).
    void action_features(const VertexStreams& streams, const float* conditions, uint32_t width, uint64_t input_faces,
                         std::vector<float>& rows);
    // The learned order of the reference's GPU action state
    // (src/neural/action_gpu.cu:1751-1798, 809-844): the policy's first
    // output on every legal action's feature row (policy.inputs() wide;
    // architecture 4 carries preserve_uv in feature 79), ranked as rank().
    void rank_policy(Policy& policy, const VertexStreams& streams, const float* conditions, uint64_t input_faces,
                     bool preserve_uv, uint32_t* order_out);
    // decode_placements / default_placements (src/neural/action_gpu.cu:
    // 1157-1200): 9 floats per legal action (position, then the unit normal
    // of each destination wedge; zeros without normals or a second wedge),
    // from a placement policy (architecture 3 or 4) on the action rows, or
    // the destination vertex without one.
    void placements(Policy* policy, const VertexStreams& streams, const float* conditions, uint64_t input_faces,
                    const PlacementOptions& options, std::vector<float>& out);
    // Wedge mapping of every legal action: wedges, then (from, to) index
    // pairs in ascending order, 0xFFFFFFFF for an absent second pair.
    void mapping(std::vector<uint32_t>& out);
    double build_ms() const { return build_ms_; }
    // max(1e-20, diameter of the input's bounds), the scale of the
    // shortest-edge and current-plane scores (src/neural/action.cpp:29).
    double diameter() const { return diameter_; }
    int device() const { return device_; }

  private:
    void build(const HostMesh& m, const float* uv);
    void claim(const uint32_t* rows, size_t n, uint32_t maximum, uint64_t removable, bool strict,
               std::vector<uint32_t>& picked);
    uint32_t claim_device(const uint32_t* drows, size_t n, uint32_t maximum, uint64_t removable, bool strict,
                          std::vector<uint32_t>& picked);
    ActionTopo topo() const;
    void upload_attributes(const VertexStreams& streams);
    void graph_features_device(const VertexStreams& streams, float* device_out);
    float* grow_vertex_features();
    // Feature rows of every legal action on the device (feature_rows_);
    // x79 < 0 keeps the target's channel 23 in feature 79.
    float* action_rows_device(const VertexStreams& streams, const float* conditions, uint32_t width,
                              uint64_t input_faces, float x79);
    void rank_scores(const float* scores, uint32_t stride, uint32_t* order_out);
    void run_rank(const void* input, uint32_t* order_out);

    int device_;
    hipStream_t stream_{};
    uint32_t vertices_ = 0, faces_ = 0, edges_ = 0, legal_count_ = 0, ranked_ = 0;
    bool has_materials_ = false;
    double build_ms_ = 0, diameter_ = 1;
    DeviceBuffer positions_, uv_, indices_, materials_;
    DeviceBuffer geometry_, invalid_, boundary_;
    DeviceBuffer offsets_, neighbors_, face_offsets_, face_lists_, face_geo_;
    DeviceBuffer edge_keys_;
    DeviceBuffer legal_;  // LegalAction per legal action
    DeviceBuffer used_, selection_, scratch_a_, scratch_b_, scratch_c_, temp_;
    DeviceBuffer order_, rank_keys_, rank_idx_, claim_out_;
    void* staging_ = nullptr;  // pinned host memory for trial downloads
    DeviceBuffer touch_;       // per-vertex lane masks of the claim kernel
    bool touch_ready_ = false;
    bool has_uv_ = false;
    // Build workspaces, kept across rebuilds.
    DeviceBuffer build_weld_, build_weld2_, build_corners_, build_temp_, build_all_, build_flags_;
    // Host positions the device weld was computed from.
    std::vector<float> weld_positions_;
    bool weld_ready_ = false;
    // Graph and action feature workspaces.
    DeviceBuffer feature_attr_, feature_work_, feature_vertex_, feature_rows_, policy_out_, placement_work_;
    const uint32_t* feature_colors_ = nullptr;
    const float *feature_normals_ = nullptr, *feature_tangents_ = nullptr;
};

// The gate of the executor: true when the candidate (index buffer, kept
// materials) is accepted.
using ActionGate = std::function<bool(const std::vector<uint32_t>&, const std::vector<uint16_t>&)>;
// The gate over the rest of a halving ladder at once: candidates in ladder
// order (each a shorter prefix of the rejected batch); returns the index of
// the first accepted one, -1 when none is. It must decide each candidate as
// the one-candidate gate would.
using ActionLadderGate =
    std::function<int(const std::vector<std::vector<uint32_t>>&, const std::vector<std::vector<uint16_t>>&)>;
struct ExecuteStats {
    uint64_t ranked = 0, trials = 0, accepted = 0, rejected = 0;
    uint32_t states = 0;  // action states built (one per accepted batch, plus the last)
    double ms = 0;
    // Wall time of each stage: state builds, ranking, batch selection,
    // trials, gate calls.
    double build_ms = 0, rank_ms = 0, select_ms = 0, trial_ms = 0, gate_ms = 0;
    // Of gate_ms: the next state's build and ranking, run while the gate
    // audits the trial (kept when it passes).
    double speculative_ms = 0, speculative_build_ms = 0;
    uint32_t accepted_batches = 0;
    // blitz::NeuralActionStop (include/blitz/neural.hpp:22-30): 0 target
    // reached, 1 trial budget, 2 no legal actions, 3 no accepted action.
    uint8_t stop_reason = 3;
};
// State kept across execute_actions calls on one input (the reference keeps
// its GpuActionState and resets it per proposal, src/neural/generate.cpp:
// 338-346): the input's action state ranked once, and an action state whose
// device buffers (and weld) are reused for every later state.
struct ExecutorCache {
    std::vector<uint32_t> input_indices;
    std::vector<uint16_t> input_materials;
    std::vector<float> input_uv;
    std::unique_ptr<ActionSet> input_state;
    int input_ranking = -1;
    std::unique_ptr<ActionSet> work;
    std::unique_ptr<ActionSet> spec;  // the speculative next state
};
// neural::execute_actions (src/neural/action.cpp:453-525) on the device
// action state: rank (0 learned: BF16 logits of `scorer` on each state's
// graph, built with `bounds`; 1 constant, 3 shortest edge, 4 current plane),
// take the greedy batch of up to `batch` actions from the rank position,
// trial it, ask the gate, commit or halve; a rejected singleton moves past
// the batch's first action. With `ladder`, once a batch is rejected the
// shorter prefixes (as many as the trial budget leaves) are trialled and
// audited in one ladder call; output and counters are the one-by-one
// loop's. Returns the final index buffer.
std::vector<uint32_t> execute_actions(int device, const HostMesh& input, const float* uv, uint64_t target,
                                      uint32_t budget, uint32_t ranking, uint32_t batch, const ActionGate& gate,
                                      ExecuteStats& stats, std::vector<uint16_t>* materials_out = nullptr,
                                      Scorer* scorer = nullptr, const Bounds* bounds = nullptr,
                                      ExecutorCache* cache = nullptr, const ActionLadderGate* ladder = nullptr);

}  // namespace blitz_hip
