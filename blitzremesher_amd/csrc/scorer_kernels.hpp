// Launch wrappers for the scorer kernels (scorer_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>

#include "scorer_types.hpp"

namespace blitz_hip {

struct FeatureInput {
    const float4* positions;
    const uint32_t* indices;
    const uint32_t* face_offsets;  // vertex -> incident faces, ascending
    const uint32_t* faces;
    const uint32_t* offsets;       // vertex -> neighbours, ascending
    const uint32_t* neighbors;
    const uint32_t* boundary;      // boundary edges per vertex
    double cx, cy, cz, diameter;
    uint32_t vertices;
    float* out;  // vertices x kFeatures
};

struct ScoreInput {
    const float* outputs;  // actions x kOutputs (learned ranking)
    const float4* positions;
    const uint32_t* indices;
    const uint32_t* face_offsets;
    const uint32_t* faces;
    const uint32_t* src;
    const uint32_t* dst;
    double scale;  // bounding-sphere diameter
    uint32_t actions;
    uint32_t mode;  // 0 learned, 1 shortest edge, 2 current plane
    float* keys;
    uint32_t* order;
};

struct DecodeInput {
    const float* outputs;
    const float4* positions;
    const uint32_t* src;
    const uint32_t* dst;
    const uint32_t* selection;  // optional action ids
    uint32_t count;
    float* out;       // count x 3
    uint32_t* error;  // set to 2 on a non-finite prediction
};

constexpr int kTeacherCandidates = 10;

struct TeacherInput {
    const float4* positions;
    const uint32_t* indices;
    const uint32_t* face_offsets;  // vertex -> incident faces, ascending
    const uint32_t* faces;
    const uint32_t* src;
    const uint32_t* dst;
    const uint32_t* selection;  // action ids
    uint32_t count;
    float* positions_out;  // count x kTeacherCandidates x 3
    float* targets_out;    // count x kTeacherCandidates x 3
};

// Offset of layer l's matrix in the BF16 copy holding only the matrices of
// the four MFMA layers.
constexpr uint32_t bf16_offset(int l) {
    uint32_t at = 0;
    for (int i = 0; i < l; ++i) at += kHidden * kLayerIn[i];
    return at;
}
constexpr uint32_t kBf16WeightCount = bf16_offset(4);

void launch_features(const FeatureInput& in, hipStream_t stream);
// h_a, h_b: vertices x kHidden; z: max(vertices x 2 kHidden, actions x kHidden).
void launch_forward_fp32(const float* x, GraphView g, const float* weights, float* h_a, float* h_b, float* z,
                         float* out, hipStream_t stream);
void launch_convert_bf16(const float* x, __bf16* y, size_t n, hipStream_t stream);
// y[i] = x[i] + add for i < n (id arrays of a graph placed after others).
void launch_offset_copy(const uint32_t* x, uint32_t* y, size_t n, uint32_t add, hipStream_t stream);
// x: vertices x kFeatures (BF16); h_a, h_b: vertices x kHidden (BF16);
// z: vertices x 2 kHidden (BF16 scratch); p: vertices x 2 kHidden (BF16 scratch).
void launch_forward_bf16(const __bf16* x, GraphView g, const __bf16* wq, const float* weights, __bf16* h_a,
                         __bf16* h_b, __bf16* z, __bf16* p, float* out, hipStream_t stream);
// Single layers, for training (activations of every layer kept).
void launch_aggregate_fp32(const float* h, GraphView g, uint32_t width, float* z, hipStream_t stream);
void launch_compensated(const float* input, const uint32_t* src, const uint32_t* dst, const float* weights,
                        const float* bias, float* output, uint32_t rows, uint32_t in, uint32_t out, bool relu,
                        hipStream_t stream);
void launch_mp_bf16(int layer, const __bf16* h, GraphView g, const __bf16* wq, const float* weights, __bf16* out,
                    __bf16* z, hipStream_t stream);
// Edge MLP over g.actions entries (src = g.src, dst = g.neighbors) in
// factorised form; p (vertices x 2 kHidden BF16) is scratch; the hidden layer
// is saved to hidden_out when it is not null.
void launch_edge_bf16(const __bf16* h, GraphView g, const __bf16* wq, const float* weights, float* out,
                      float* hidden_out, __bf16* p, hipStream_t stream);
void launch_scores(const ScoreInput& in, hipStream_t stream);
size_t sort_workspace(uint32_t n);
hipError_t launch_sort(void* temp, size_t temp_bytes, const float* keys_in, float* keys_out, const uint32_t* order_in,
                       uint32_t* order_out, uint32_t n, hipStream_t stream);
void launch_decode(const DecodeInput& in, hipStream_t stream);
void launch_teacher(const TeacherInput& in, hipStream_t stream);

}  // namespace blitz_hip
