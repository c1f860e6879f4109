// Launch wrappers for the scorer training kernels (train_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>

#include "scorer_types.hpp"

namespace blitz_hip {

struct LossInput {
    uint32_t count;     // batch entries
    uint32_t pool;      // entries per ranking pool (<= 256)
    uint32_t pairs;     // sum over pools of preferred * (valid - preferred)
    uint32_t valid;     // valid entries
    uint32_t placed;    // entries with a placement target
    float margin, penalty, placement_weight;
};

struct AdamStep {
    float clip, lr, decay, correction1, correction2;
};

// Loss gradient dy and, in *loss, the loss: per-wave partials (`parts`,
// loss_parts(in) floats) added in a fixed order.
uint32_t loss_parts(const LossInput& in);
void launch_loss(const float* y, const uint8_t* labels, const float* targets, float* dy, float* parts, float* loss,
                 const LossInput& in, hipStream_t stream);
// c[m][n] += sum over rows of (a masked)[r][m] * b[r][n] and, when `sums` is
// set, sums[m] += sum over rows of (a masked)[r][m]. Row slices write partial
// tiles into `work` (gemm_at_b_workspace floats), added in slice order: the
// same bits on every run.
size_t gemm_at_b_workspace(uint32_t rows, uint32_t m, uint32_t n);
void launch_gemm_at_b(const float* a, uint32_t lda, const float* a_mask, const float* b, uint32_t ldb, float* c,
                      float* sums, float* work, uint32_t rows, uint32_t m, uint32_t n, hipStream_t stream);
// t[r][n] = sum_k (a masked)[r][k] * w[k][n].
void launch_gemm_a_w(const float* a, const float* a_mask, const float* w, float* t, uint32_t rows, uint32_t k,
                     uint32_t n, hipStream_t stream);
void launch_batch_endpoints(const uint32_t* actions, GraphView g, uint32_t* src, uint32_t* dst, uint32_t count,
                            hipStream_t stream);
void launch_gather_edge(const float* h, const uint32_t* src, const uint32_t* dst, float* z, uint32_t count,
                        hipStream_t stream);
// dh[v] = sum, in batch order, of the edge-input gradients dz of the batch
// entries that read vertex v (the transpose of gather_edge), every vertex
// written. Workspace: counts and offsets vertices + 1, cursor vertices, list
// 2 * count entries, and gather_edge_scan_bytes(vertices) of scan scratch.
size_t gather_edge_scan_bytes(uint32_t vertices);
void launch_gather_edge_backward(const float* dz, const uint32_t* src, const uint32_t* dst, uint32_t count,
                                 uint32_t vertices, uint32_t* counts, uint32_t* offsets, uint32_t* cursor,
                                 uint32_t* list, void* scan_temp, size_t scan_bytes, float* dh, hipStream_t stream);
void launch_aggregate_backward(const float* t, GraphView g, uint32_t width, float* dh, hipStream_t stream);
// *sum = sum of g[i]^2 (non-finite entries give +inf), per-wave partials in
// `parts` (4096 floats) added in a fixed order.
void launch_norm(const float* g, uint32_t n, float* parts, float* sum, hipStream_t stream);
void launch_adamw(float* value, const float* grad, float* mean, float* variance, uint32_t n, const AdamStep& s,
                  hipStream_t stream);
void launch_bf16_to_fp32(const __bf16* x, float* y, size_t n, hipStream_t stream);

}  // namespace blitz_hip
