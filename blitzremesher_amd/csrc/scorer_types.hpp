// Plain data shared by the scorer host controller and kernels.
#pragma once
#include <cstdint>

namespace blitz_hip {

// Edge-graph scorer dimensions: 32 vertex features, three mean-aggregation
// message-passing layers of width 128, an edge MLP over [h_src, h_dst]
// (256 -> 128) and a 4-wide head (logit, placement xyz).
constexpr uint32_t kFeatures = 32;
constexpr uint32_t kHidden = 128;
constexpr uint32_t kLayers = 3;
constexpr uint32_t kEdgeIn = 2 * kHidden;
constexpr uint32_t kOutputs = 4;

// Row-major [out, in] weights followed by the bias, layer after layer (the
// reference's portable float32 layout, src/neural/internal.hpp:42-43).
constexpr uint32_t kLayerIn[5] = {2 * kFeatures, 2 * kHidden, 2 * kHidden, kEdgeIn, kHidden};
constexpr uint32_t kLayerOut[5] = {kHidden, kHidden, kHidden, kHidden, kOutputs};
constexpr uint32_t layer_offset(int l) {
    uint32_t at = 0;
    for (int i = 0; i < l; ++i) at += kLayerOut[i] * (kLayerIn[i] + 1);
    return at;
}
constexpr uint32_t kWeightCount = layer_offset(5);

// Graph of one mesh: undirected neighbour CSR (sorted rows). Every CSR entry
// a is one directed contraction action src[a] -> neighbors[a]; reverse[a] is
// the entry of the opposite action.
struct GraphView {
    const uint32_t* offsets;    // vertices + 1
    const uint32_t* neighbors;  // actions
    const uint32_t* src;        // actions
    const uint32_t* reverse;    // actions
    uint32_t vertices, actions;
};

}  // namespace blitz_hip
