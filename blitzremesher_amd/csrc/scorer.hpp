// Host controller of the edge-graph proposal scorer: mesh topology (host),
// vertex features, FP32/BF16 forward passes, ranking and placement decode
// (HIP). Counterpart of the reference's ActionCuda / GpuActionState ranking
// (src/neural/action.hpp:75, src/neural/action_gpu.cu:1751-1798).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "audit.hpp"
#include "scorer_types.hpp"

namespace blitz_hip {

// Host topology of one mesh (see scorer_types.hpp GraphView).
struct Topology {
    std::vector<uint32_t> offsets, neighbors, src, reverse;
    std::vector<uint32_t> face_offsets, faces;  // vertex -> incident face corners, ascending face
    std::vector<uint32_t> boundary;             // edges with one incident face, per vertex
};
Topology build_topology(const uint32_t* indices, size_t faces, size_t vertices);

struct DeviceBuffer {
    void* p = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer();
    void resize(size_t n);
    template <class T> T* as() const { return static_cast<T*>(p); }
};

struct Graph {
    uint32_t vertices = 0, actions = 0, faces = 0;
    double diameter = 0;
    DeviceBuffer positions, indices, offsets, neighbors, src, reverse, face_offsets, faces_of, boundary;
    DeviceBuffer features, features_bf16, outputs, h_a, h_b, z, hq_a, hq_b;
    DeviceBuffer keys, order, keys_out, order_out, sort_temp;
    bool has_outputs = false;
    GraphView view() const;
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
    void forward(Graph& g, Precision p);
    void download_outputs(const Graph& g, float* out);
    void download_features(const Graph& g, float* out);
    void rank(Graph& g, Ranking r, uint32_t* order);
    void decode(Graph& g, const uint32_t* selection, size_t count, double* out);
    double last_forward_ms() const { return forward_ms_; }
    int device() const { return device_; }
    hipStream_t stream() const { return stream_; }

  private:
    int device_;
    hipStream_t stream_{};
    hipEvent_t begin_{}, end_{};
    DeviceBuffer weights_, weights_bf16_;
    double forward_ms_ = 0;
    uint32_t blocks_ = 512;
};

}  // namespace blitz_hip
