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

void launch_loss(const float* y, const uint8_t* labels, const float* targets, float* dy, float* loss,
                 const LossInput& in, hipStream_t stream);
// c[m][n] += sum over rows of (a masked)[r][m] * b[r][n] and, when `sums` is
// set, sums[m] += sum over rows of (a masked)[r][m].
void launch_gemm_at_b(const float* a, uint32_t lda, const float* a_mask, const float* b, uint32_t ldb, float* c,
                      float* sums, uint32_t rows, uint32_t m, uint32_t n, hipStream_t stream);
// t[r][n] = sum_k (a masked)[r][k] * w[k][n].
void launch_gemm_a_w(const float* a, const float* a_mask, const float* w, float* t, uint32_t rows, uint32_t k,
                     uint32_t n, hipStream_t stream);
void launch_batch_endpoints(const uint32_t* actions, GraphView g, uint32_t* src, uint32_t* dst, uint32_t count,
                            hipStream_t stream);
void launch_gather_edge(const float* h, const uint32_t* src, const uint32_t* dst, float* z, uint32_t count,
                        hipStream_t stream);
void launch_scatter_edge(const float* dz, const uint32_t* src, const uint32_t* dst, float* dh, uint32_t count,
                         hipStream_t stream);
void launch_aggregate_backward(const float* t, GraphView g, uint32_t width, float* dh, hipStream_t stream);
void launch_norm(const float* g, uint32_t n, float* sum, hipStream_t stream);
void launch_adamw(float* value, const float* grad, float* mean, float* variance, uint32_t n, const AdamStep& s,
                  hipStream_t stream);
void launch_bf16_to_fp32(const __bf16* x, float* y, size_t n, hipStream_t stream);

}  // namespace blitz_hip
