// Host controller of the edge-graph proposal scorer: mesh topology,
// vertex features, FP32/BF16 forward passes, ranking and placement decode
// (HIP). Counterpart of the reference's ActionCuda / GpuActionState ranking
// (src/neural/action.hpp:75, src/neural/action_gpu.cu:1751-1798).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "audit.hpp"
#include "device_buffer.hpp"
#include "scorer_types.hpp"

namespace blitz_hip {



struct Graph {
    uint32_t vertices = 0, actions = 0, faces = 0;
    uint32_t corners = 0;  // entries of faces_of
    double diameter = 0;   // of the bounds the features were computed with; 0 for a union of graphs
    DeviceBuffer positions, indices, offsets, neighbors, src, reverse, face_offsets, faces_of, boundary;
    DeviceBuffer features, features_bf16, outputs, h_a, h_b, z, hq_a, hq_b, zq, pq;
    DeviceBuffer keys, order, keys_out, order_out, sort_temp;
    // Scratch of the device topology build.
    DeviceBuffer build_keys, build_sorted, build_unique, build_counts, build_values, build_temp;
    // Training: activations of every layer (FP32 and BF16), layer inputs,
    // batch buffers and gradients.
    DeviceBuffer t_h[4], t_hq[4], t_z[3];
    DeviceBuffer batch_actions, batch_src, batch_dst, labels, targets, hidden, batch_out, dy, t_e, z_e, dz_e, dh_a, dh_b, t_v,
        accum, reduce_work, edge_buckets;
    bool has_outputs = false;
    GraphView view() const;
};

// Loss and optimizer settings of one training step (defaults of the
// reference's UpdateSettings, training/update_cuda.hpp:47-53).
struct TrainSettings {
    float margin = 1.f, penalty = 1e-4f, lr = 1e-3f, decay = 1e-4f, max_norm = 1.f, placement_weight = 1.f;
    uint32_t pool = 32;
    uint8_t precision = 1;  // 0 FP32 forward, 1 BF16 forward
};
struct TrainResult {
    float loss = 0, grad_norm = 0;
    bool applied = false;  // false when the gradient norm is not finite
};

enum class Precision : uint8_t { Fp32 = 0, Bf16 = 1 };
// Values of the reference's NeuralRanking (include/blitz/neural.hpp:184).
enum class Ranking : uint8_t { Learned = 0, Constant = 1, Shuffled = 2, ShortestEdge = 3, CurrentPlane = 4 };

class Scorer {
  public:
    Scorer(int device, const float* weights, size_t count);
    ~Scorer();
    Scorer(const Scorer&) = delete;
    Scorer& operator=(const Scorer&) = delete;
    void set_weights(const float* weights, size_t count);
    std::unique_ptr<Graph> graph(const HostMesh& m, const Bounds& b);
    // graph() into an existing graph, reusing its device buffers where they
    // are large enough (one graph per mesh state of an executor run). With
    // `same_positions` the graph keeps the positions of its previous build
    // (the caller guarantees the vertex stream is unchanged).
    void rebuild(Graph& g, const HostMesh& m, const Bounds& b, bool same_positions = false);
    // Disjoint union of graphs built by this scorer, in order: the vertex,
    // action, face and face-list ids of each part are offset by the counts
    // of the parts before it, and each part keeps the vertex features of its
    // own bounds, so one training batch or forward pass spans many meshes.
    std::unique_ptr<Graph> concat(const Graph* const* parts, size_t count);
    void forward(Graph& g, Precision p);
    void download_outputs(const Graph& g, float* out);
    void download_features(const Graph& g, float* out);
    void rank(Graph& g, Ranking r, uint32_t* order);
    void decode(Graph& g, const uint32_t* selection, size_t count, float* out);
    // Ten teacher positions per action and their clamped offset targets.
    void teacher(Graph& g, const uint32_t* selection, size_t count, float* positions, float* targets);
    double last_forward_ms() const { return forward_ms_; }
    int device() const { return device_; }
    hipStream_t stream() const { return stream_; }
    // One optimizer step on a batch of labelled actions of graph g (see
    // train_kernels.hip for the loss). labels: bit 0 valid, bit 1 preferred,
    // bit 2 placement target present; targets: 3 floats per entry or null.
    TrainResult train_step(Graph& g, const uint32_t* actions, const uint8_t* labels, const float* targets,
                           size_t count, const TrainSettings& s);
    void download_gradients(float* out);
    void download_weights(float* out);
    void reset_optimizer();
    double last_step_ms() const { return step_ms_; }
    uint32_t optimizer_steps() const { return step_; }

  private:
    void refresh_bf16();
    int device_;
    hipStream_t stream_{};
    hipEvent_t begin_{}, end_{};
    DeviceBuffer weights_, weights_bf16_;
    DeviceBuffer grads_, mean_, variance_;
    uint32_t step_ = 0;
    double step_ms_ = 0;
    double forward_ms_ = 0;
    uint32_t blocks_ = 512;
};

}  // namespace blitz_hip
