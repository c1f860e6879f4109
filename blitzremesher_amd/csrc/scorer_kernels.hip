// Edge-graph proposal scorer kernels for gfx950.
//
// Network (the build's scorer; the reference's production policy is an
// independent-action MLP, src/neural/action.hpp:6-10, and its legacy encoder a
// 24-feature, width-64 graph network, src/neural/inference.cu:117-150):
//   h0 = 32 vertex features;
//   h(l+1) = ReLU(W_l [h_l ; mean_{w in N(v)} h_l(w)] + b_l), l = 0..2, width 128
//            (mean aggregation as src/neural/inference.cu:18-30);
//   for each directed edge a = (u -> v): g = ReLU(W_e [h3(u) ; h3(v)] + b_e),
//   out(a) = W_o g + b_o: logit and a clamped placement offset.
//
// Two numeric modes:
//  * FP32: deterministic. Aggregation sums neighbours in CSR order; every
//    linear layer is the reference's compensated kernel
//    (src/neural/inference.cu:75-108): eight strided Kahan partials per output
//    combined in fixed order with the bias. A CPU restatement reproduces it
//    bit for bit, so FP32 rankings are reproducible.
//  * BF16: the fast path. Activations and weights in BF16, products on the
//    gfx950 matrix cores (v_mfma_f32_16x16x32_bf16) with FP32 accumulation.
//    Each message-passing layer is a high-occupancy aggregation kernel
//    (neighbour sums in FP32) into a BF16 [h ; mean] buffer followed by a
//    persistent MFMA GEMM with a bias/ReLU epilogue. The edge layer is
//    factorised: W_e [h_u ; h_v] = W_e(:, :128) h_u + W_e(:, 128:) h_v, so one
//    GEMM over the vertices gives both halves and each action only gathers
//    two 128-vectors, adds, applies ReLU and the 128 -> 4 head in FP32.
//
// Vertex features (32, FP64 internally, stored FP32); p = (P - centre) / diameter:
//   0-2 p | 3-5 unit area-weighted normal | 6-15 area-normalised plane quadric
//   (a00 a01 a02 a03 a11 a12 a13 a22 a23 a33, the Quadric layout of
//   src/simplify.cpp:10-32) | 16 curvature: n . (mean neighbour p - p) / mean
//   edge, clamped to [-1,1] | 17 valence / 16 | 18 on boundary | 19 UV stream
//   present | 20 min(64 * mean edge, 4) | 21 incident faces / 8 | 22 boundary
//   edges / 4 | 23-25 linear colour (white) | 26-27 uv / 16 | 28 normal stream
//   present | 29 colour stream present | 30 min(sqrt(area) / mean edge, 4) |
//   31 constant 1.
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "scorer_kernels.hpp"
#include "device_sort.hpp"
#include "vec_device.hpp"

namespace blitz_hip {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ features

__device__ __forceinline__ void normalised(const FeatureInput& in, uint32_t i, double* p) {
    const float4 q = in.positions[i];
    p[0] = (double(q.x) - in.cx) / in.diameter;
    p[1] = (double(q.y) - in.cy) / in.diameter;
    p[2] = (double(q.z) - in.cz) / in.diameter;
}

__global__ __launch_bounds__(256) void features_kernel(FeatureInput in) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= in.vertices) return;
    double p[3];
    normalised(in, v, p);
    double normal[3] = {0, 0, 0}, quad[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, total = 0;
    const uint32_t fb = in.face_offsets[v], fe = in.face_offsets[v + 1];
    for (uint32_t k = fb; k < fe; ++k) {
        const uint32_t f = in.faces[k];
        double a[3], b[3], c[3];
        normalised(in, in.indices[3 * f], a);
        normalised(in, in.indices[3 * f + 1], b);
        normalised(in, in.indices[3 * f + 2], c);
        const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double w[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const double cr[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
        const double area = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        double plane[4] = {0, 0, 0, 0};
        if (area > 0) {
            plane[0] = cr[0] / area;
            plane[1] = cr[1] / area;
            plane[2] = cr[2] / area;
        }
        plane[3] = -(plane[0] * a[0] + plane[1] * a[1] + plane[2] * a[2]);
        int k2 = 0;
        for (int i = 0; i < 4; ++i)
            for (int j = i; j < 4; ++j) quad[k2++] += plane[i] * plane[j] * area;
        normal[0] += cr[0];
        normal[1] += cr[1];
        normal[2] += cr[2];
        total += area;
    }
    const double nl = sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
    double unit[3] = {0, 0, 0};
    if (nl > 0) {
        unit[0] = normal[0] / nl;
        unit[1] = normal[1] / nl;
        unit[2] = normal[2] / nl;
    }
    const uint32_t nb = in.offsets[v], ne = in.offsets[v + 1];
    double sum[3] = {0, 0, 0}, edge_sum = 0;
    for (uint32_t k = nb; k < ne; ++k) {
        double q[3];
        normalised(in, in.neighbors[k], q);
        sum[0] += q[0];
        sum[1] += q[1];
        sum[2] += q[2];
        const double d[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]};
        edge_sum += sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    }
    const uint32_t deg = ne - nb;
    const double degc = double(deg > 0 ? deg : 1u);
    const double mean_edge = edge_sum / degc;
    const double lap[3] = {sum[0] / degc - p[0], sum[1] / degc - p[1], sum[2] / degc - p[2]};
    const double curvature =
        fmin(fmax((unit[0] * lap[0] + unit[1] * lap[1] + unit[2] * lap[2]) / fmax(mean_edge, 1e-30), -1.0), 1.0);
    const uint32_t bcount = in.boundary[v];
    const double denom = fmax(total, 1e-30);
    float* x = in.out + size_t(v) * kFeatures;
    x[0] = float(p[0]);
    x[1] = float(p[1]);
    x[2] = float(p[2]);
    x[3] = float(unit[0]);
    x[4] = float(unit[1]);
    x[5] = float(unit[2]);
    for (int i = 0; i < 10; ++i) x[6 + i] = float(quad[i] / denom);
    x[16] = float(curvature);
    x[17] = float(double(deg) / 16.0);
    x[18] = bcount ? 1.f : 0.f;
    x[19] = 0.f;
    x[20] = float(fmin(mean_edge * 64.0, 4.0));
    x[21] = float(double(fe - fb) / 8.0);
    x[22] = float(double(bcount) / 4.0);
    x[23] = x[24] = x[25] = 1.f;
    x[26] = x[27] = 0.f;
    x[28] = x[29] = 0.f;
    x[30] = float(fmin(sqrt(total) / fmax(mean_edge, 1e-30), 4.0));
    x[31] = 1.f;
}

// ------------------------------------------------------------ FP32 (exact)

// z[v] = [h[v], mean of h over N(v)], neighbours summed in CSR order. Four
// channels per thread (width is a multiple of 4); neighbour indices are read
// four at a time so their rows' loads are in flight together, and the adds
// still run in CSR order, as with one channel per thread.
__device__ __forceinline__ float4 add4(float4 a, float4 b) {
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

__device__ __forceinline__ float4 row4(const float* __restrict__ p, size_t at) {
    return *reinterpret_cast<const float4*>(p + at);
}

__global__ __launch_bounds__(256) void aggregate_fp32(const float* __restrict__ h, GraphView g, uint32_t width,
                                                      float* __restrict__ z) {
    const uint32_t quads = width / 4;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(g.vertices) * quads) return;
    const uint32_t v = uint32_t(i / quads), c = uint32_t(i % quads) * 4;
    const uint32_t b = g.offsets[v], e = g.offsets[v + 1];
    float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t k = b;
    for (; k + 4 <= e; k += 4) {
        const uint32_t n0 = g.neighbors[k], n1 = g.neighbors[k + 1], n2 = g.neighbors[k + 2],
                       n3 = g.neighbors[k + 3];
        const float4 a0 = row4(h, size_t(n0) * width + c), a1 = row4(h, size_t(n1) * width + c),
                     a2 = row4(h, size_t(n2) * width + c), a3 = row4(h, size_t(n3) * width + c);
        sum = add4(add4(add4(add4(sum, a0), a1), a2), a3);
    }
    for (; k < e; ++k) sum = add4(sum, row4(h, size_t(g.neighbors[k]) * width + c));
    const float d = float(e - b > 0 ? e - b : 1u);
    *reinterpret_cast<float4*>(z + size_t(v) * 2 * width + c) = row4(h, size_t(v) * width + c);
    *reinterpret_cast<float4*>(z + size_t(v) * 2 * width + width + c) =
        make_float4(sum.x / d, sum.y / d, sum.z / d, sum.w / d);
}

// Compensated linear layer (src/neural/inference.cu:75-108). Eight lanes per
// G outputs of one row (G = 4 or 8): lane j takes k = j, j + 8, ... and keeps
// one compensated (Kahan) sum per output, in the order eight lanes per output
// would, so each output has the same bits; the input value is loaded once for
// the G outputs, whose sums are independent dependency chains. Output q's
// eight partial sums and residuals are then folded in lane order by lane q,
// the G folds running side by side. A block takes one group of G outputs and
// kCompRows rows per eight-lane group; the G weight rows are staged in LDS
// once per block. With `src` set, row r's input is [h[src[r]] ; h[dst[r]]].
// out % G == 0, in <= kCompMaxIn (G = 1 takes any output count).
constexpr uint32_t kCompMaxIn = 256, kCompRows = 4, kCompRowsPerBlock = 32 * kCompRows;
static_assert(kEdgeIn <= kCompMaxIn && 2 * kHidden <= kCompMaxIn && 2 * kFeatures <= kCompMaxIn,
              "layer inputs must fit the staged weight rows");

template <int G>
__global__ __launch_bounds__(256) void compensated_linear(const float* __restrict__ input, const uint32_t* src,
                                                          const uint32_t* dst, const float* __restrict__ weights,
                                                          const float* __restrict__ bias, float* __restrict__ output,
                                                          uint32_t rows, uint32_t in, uint32_t out, bool relu) {
    __shared__ float w[G * kCompMaxIn];
    const uint32_t groups = out / G;
    const uint32_t c0 = (blockIdx.x % groups) * G;
    for (uint32_t e = threadIdx.x; e < G * in; e += blockDim.x) w[e] = weights[size_t(c0) * in + e];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 7;
    const uint32_t split = src ? in / 2 : in;
    for (uint32_t rr = 0; rr < kCompRows; ++rr) {
        const size_t row = size_t(blockIdx.x / groups) * kCompRowsPerBlock + rr * 32 + threadIdx.x / 8;
        const bool live = row < rows;
        // Input of this row: k < split from x0[k], the rest from x1[k - split].
        const float* x0 = input + (live ? row : 0) * in;
        const float* x1 = x0;
        if (src && live) {
            x0 = input + size_t(src[row]) * split;
            x1 = input + size_t(dst[row]) * split;
        }
        float sum[G], correction[G];
#pragma unroll
        for (int q = 0; q < G; ++q) sum[q] = correction[q] = 0.f;
        if (live)
            for (uint32_t k = lane; k < in; k += 8) {
                const float x = k < split ? x0[k] : x1[k - split];
#pragma unroll
                for (int q = 0; q < G; ++q) {
                    const float term = __fmul_rn(x, w[q * in + k]) - correction[q];
                    const float next = __fadd_rn(sum[q], term);
                    correction[q] = __fsub_rn(__fsub_rn(next, sum[q]), term);
                    sum[q] = next;
                }
            }
        const bool folds = live && lane < G;
        float total = folds ? bias[c0 + lane] : 0.f, error = 0.f;
        for (int j = 0; j < 8; ++j) {
            float part = 0.f, residual = 0.f;
#pragma unroll
            for (int q = 0; q < G; ++q) {
                const float p = __shfl(sum[q], j, 8), r = __shfl(correction[q], j, 8);
                if (lane == uint32_t(q)) {
                    part = p;
                    residual = r;
                }
            }
            if (folds) {
                float term = __fsub_rn(part, error), next = __fadd_rn(total, term);
                error = __fsub_rn(__fsub_rn(next, total), term);
                total = next;
                term = __fsub_rn(-residual, error);
                next = __fadd_rn(total, term);
                error = __fsub_rn(__fsub_rn(next, total), term);
                total = next;
            }
        }
        if (folds) output[row * out + c0 + lane] = relu && total < 0 ? 0.f : total;
    }
}

// Grid of compensated_linear<G>: one block per output group and
// kCompRowsPerBlock rows.
inline dim3 comp_grid(uint32_t rows, uint32_t out, uint32_t g) {
    return dim3((out / g) * ((rows + kCompRowsPerBlock - 1) / kCompRowsPerBlock));
}

// ------------------------------------------------------------- BF16 / MFMA

// v_mfma_f32_16x16x32_bf16 operand layout (wave64): lane l supplies A row
// (l % 16), k = 8 * (l / 16) .. +7, and B column (l % 16), same k; it receives
// D rows 4 * (l / 16) .. +3 of column (l % 16).
constexpr uint32_t kTileRows = 64;
constexpr uint32_t kPad = 8;  // bf16 elements of row padding in LDS

__device__ __forceinline__ floatx4 mfma(bf16x8 a, bf16x8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// z[v] = [h[v] ; mean of h over N(v)] in BF16 (K = 2 * width): one thread per
// 8 channels of one vertex, neighbour sums in FP32 in CSR order.
template <uint32_t W>
__global__ __launch_bounds__(256) void aggregate_bf16(const __bf16* __restrict__ h, GraphView g,
                                                      __bf16* __restrict__ z) {
    constexpr uint32_t groups = W / 8;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(g.vertices) * groups) return;
    const uint32_t v = uint32_t(i / groups), c = uint32_t(i % groups) * 8;
    const bf16x8 self = *reinterpret_cast<const bf16x8*>(h + size_t(v) * W + c);
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t b = g.offsets[v], e = g.offsets[v + 1];
    uint32_t k = b;
    // Four neighbours' loads in flight together; the adds keep CSR order.
    for (; k + 4 <= e; k += 4) {
        bf16x8 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = *reinterpret_cast<const bf16x8*>(h + size_t(g.neighbors[k + u]) * W + c);
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[j] += float(q[u][j]);
    }
    for (; k < e; ++k) {
        const bf16x8 q = *reinterpret_cast<const bf16x8*>(h + size_t(g.neighbors[k]) * W + c);
        for (int j = 0; j < 8; ++j) acc[j] += float(q[j]);
    }
    const float inv = 1.f / float(e - b > 0 ? e - b : 1u);
    bf16x8 mean;
    for (int j = 0; j < 8; ++j) mean[j] = __bf16(acc[j] * inv);
    *reinterpret_cast<bf16x8*>(z + size_t(v) * 2 * W + c) = self;
    *reinterpret_cast<bf16x8*>(z + size_t(v) * 2 * W + W + c) = mean;
}

// C[:, 128 y .. 128 y + 127] = act(A B_y^T + bias) for a row-major BF16
// A [M][K] and B_y = rows 0..127 of W + y * block_offset (row stride ldw):
// persistent blocks keep B_y in LDS; each wave owns 16 rows of a 64-row tile
// and eight 16-column MFMA blocks; the A fragments of the next tile are
// loaded from HBM while the current tile is multiplied. FP32 accumulation.
template <uint32_t K>
__global__ __launch_bounds__(256) void gemm_bf16(const __bf16* __restrict__ a, uint32_t m,
                                                 const __bf16* __restrict__ wq, uint32_t ldw, uint32_t block_offset,
                                                 const float* __restrict__ bias, bool relu, __bf16* __restrict__ c,
                                                 uint32_t ldc) {
    constexpr uint32_t steps = K / 32;
    extern __shared__ __bf16 smem[];
    const __bf16* src = wq + size_t(blockIdx.y) * block_offset;
    for (uint32_t i = threadIdx.x; i < kHidden * (K / 8); i += blockDim.x) {
        const uint32_t row = i / (K / 8), col = (i % (K / 8)) * 8;
        *reinterpret_cast<bf16x8*>(smem + row * (K + kPad) + col) =
            *reinterpret_cast<const bf16x8*>(src + size_t(row) * ldw + col);
    }
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t rl = wave * 16 + (lane & 15), kq = (lane >> 4) * 8;
    const uint32_t tiles = (m + kTileRows - 1) / kTileRows;
    bf16x8 next[steps];
    auto load_a = [&](uint32_t tile) {
        const uint32_t row = tile * kTileRows + rl;
#pragma unroll
        for (uint32_t s = 0; s < steps; ++s)
            next[s] = row < m ? *reinterpret_cast<const bf16x8*>(a + size_t(row) * K + s * 32 + kq) : bf16x8{};
    };
    if (blockIdx.x < tiles) load_a(blockIdx.x);
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        bf16x8 cur[steps];
#pragma unroll
        for (uint32_t s = 0; s < steps; ++s) cur[s] = next[s];
        if (tile + gridDim.x < tiles) load_a(tile + gridDim.x);
        floatx4 acc[8];
#pragma unroll
        for (int nb = 0; nb < 8; ++nb) acc[nb] = floatx4{0, 0, 0, 0};
#pragma unroll
        for (uint32_t s = 0; s < steps; ++s) {
#pragma unroll
            for (int nb = 0; nb < 8; ++nb) {
                const bf16x8 b =
                    *reinterpret_cast<const bf16x8*>(smem + (nb * 16 + (lane & 15)) * (K + kPad) + s * 32 + kq);
                acc[nb] = mfma(cur[s], b, acc[nb]);
            }
            // Keeps the next step's B-fragment reads below this step's
            // MFMAs: hoisting all K/4 of them takes two waves' registers.
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int nb = 0; nb < 8; ++nb) {
            const uint32_t col = nb * 16 + (lane & 15);
            const float bv = bias ? bias[col] : 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t row = tile * kTileRows + wave * 16 + (lane >> 4) * 4 + i;
                float v = acc[nb][i] + bv;
                if (relu) v = fmaxf(v, 0.f);
                if (row < m) c[size_t(row) * ldc + blockIdx.y * kHidden + col] = __bf16(v);
            }
        }
    }
}

// Edge MLP in factorised form: with p = h3 [W_e(:, :128) | W_e(:, 128:)]^T
// precomputed per vertex, g(a) = ReLU(p_src(a) + p_dst(a) + b_e) and
// out(a) = W_o g + b_o (FP32). Sixteen lanes per action, eight channels each.
__global__ __launch_bounds__(256) void edge_bf16(const __bf16* __restrict__ p, GraphView g,
                                                 const float* __restrict__ be, const float* __restrict__ wo,
                                                 const float* __restrict__ bo, float* __restrict__ out,
                                                 float* __restrict__ hidden_out) {
    const uint32_t lane = threadIdx.x & 15, c = lane * 8;
    const uint32_t a = (blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const bool live = a < g.actions;
    float part[kOutputs] = {0, 0, 0, 0};
    if (live) {
        const uint32_t u = g.src[a], v = g.neighbors[a];
        const bf16x8 pa = *reinterpret_cast<const bf16x8*>(p + size_t(u) * 2 * kHidden + c);
        const bf16x8 pb = *reinterpret_cast<const bf16x8*>(p + size_t(v) * 2 * kHidden + kHidden + c);
        float hid[8];
        for (int j = 0; j < 8; ++j) hid[j] = fmaxf(float(pa[j]) + float(pb[j]) + be[c + j], 0.f);
        if (hidden_out)
            for (int j = 0; j < 8; ++j) hidden_out[size_t(a) * kHidden + c + j] = hid[j];
        for (uint32_t o = 0; o < kOutputs; ++o)
            for (int j = 0; j < 8; ++j) part[o] += hid[j] * wo[o * kHidden + c + j];
    }
    for (uint32_t o = 0; o < kOutputs; ++o)
        for (int off = 8; off > 0; off >>= 1) part[o] += __shfl_xor(part[o], off);
    if (live && lane == 0)
        for (uint32_t o = 0; o < kOutputs; ++o) out[size_t(a) * kOutputs + o] = part[o] + bo[o];
}

__global__ void to_bf16(const float* __restrict__ x, __bf16* __restrict__ y, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        y[i] = __bf16(x[i]);
}

__global__ void offset_copy(const uint32_t* __restrict__ x, uint32_t* __restrict__ y, size_t n, uint32_t add) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        y[i] = x[i] + add;
}

// ----------------------------------------------------------------- ranking

// Scores of src/neural/action_gpu.cu:809-844: learned logit, shortest edge
// (-|dst - src| / scale) or current plane (-sum over faces around src of
// area-weighted squared plane distance of dst, FP64).
__global__ void score_kernel(ScoreInput in) {
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= in.actions) return;
    float score = 0.f;
    if (in.mode == 0) {
        score = in.outputs[size_t(a) * kOutputs];
    } else if (in.mode == 3) {
        score = 0.f;  // constant ranking: action order
    } else {
        // The reference's Vec3 arithmetic (float differences and cross
        // products, double dot products), as the action state ranks.
        const double scale = fmax(1e-20, in.scale);
        const uint32_t u = in.src[a], v = in.dst[a];
        score = in.mode == 1 ? shortest_edge_score(in.positions, u, v, scale)
                             : current_plane_score(in.positions, in.indices, in.face_offsets, in.faces, u, v, scale);
    }
    // -0.0 sorts as +0.0 so equal scores keep ascending action order.
    in.keys[a] = score + 0.f;
    in.order[a] = a;
}

// Placement of src/neural/action_gpu.cu:1157-1188 without grid snapping:
// midpoint + clamp(offset, -1, 1) * edge length, FP64.
// Vec3 arithmetic of the reference's action kernels
// (src/neural/action_gpu.cu:16-38): float components, mul3 scaling in double
// with a float result, dot3/len3 in double.
__device__ __forceinline__ float3 add3(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 mul3(float3 a, double s) {
    return make_float3(float(a.x * s), float(a.y * s), float(a.z * s));
}
__device__ __forceinline__ double dot3(float3 a, float3 b) {
    return double(a.x) * b.x + double(a.y) * b.y + double(a.z) * b.z;
}
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double len3(float3 a) { return sqrt(dot3(a, a)); }
__device__ __forceinline__ float3 norm3(float3 a) {
    const double n = len3(a);
    return n > 0 ? mul3(a, 1 / n) : make_float3(0.f, 0.f, 0.f);
}
__device__ __forceinline__ float3 xyz(float4 p) { return make_float3(p.x, p.y, p.z); }
__device__ __forceinline__ float clamp1(float v) { return fminf(1.f, fmaxf(-1.f, v)); }

// decode_placements (src/neural/action_gpu.cu:1157-1188) without grid
// snapping: middle + clamp(y, -1, 1) * edge length, float positions; a
// non-finite prediction of a selected action sets the error flag.
__global__ void decode_kernel(DecodeInput in) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= in.count) return;
    const uint32_t a = in.selection ? in.selection[i] : i;
    const float* y = in.outputs + size_t(a) * kOutputs;
    for (uint32_t j = 0; j < kOutputs; ++j)
        if (!isfinite(y[j])) {
            atomicExch(in.error, 2u);
            return;
        }
    const float3 p = xyz(in.positions[in.src[a]]), q = xyz(in.positions[in.dst[a]]);
    const float3 middle = mul3(add3(p, q), .5);
    const double length = len3(sub3(p, q));
    const float3 pos = add3(middle, mul3(make_float3(clamp1(y[1]), clamp1(y[2]), clamp1(y[3])), length));
    in.out[3 * size_t(i)] = pos.x;
    in.out[3 * size_t(i) + 1] = pos.y;
    in.out[3 * size_t(i) + 2] = pos.z;
}

// teacher_placements (src/neural/action_gpu.cu:1304-1373) for position
// candidates only (no grid snapping, no normals): 0 from, 1 to, 2 middle,
// 3 the optimum of the area-weighted plane quadrics of the faces around both
// endpoints (FP64, Gauss-Jordan with partial pivoting; middle when singular),
// 4..9 middle +/- a quarter edge along x, y, z; targets are the clamped
// offsets from the middle in edge lengths.
__global__ void teacher_kernel(TeacherInput in) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= in.count) return;
    const uint32_t a = in.selection[i];
    const uint32_t from = in.src[a], to = in.dst[a];
    const float3 pf = xyz(in.positions[from]), pt = xyz(in.positions[to]);
    const float3 middle = mul3(add3(pf, pt), .5);
    const double scale = len3(sub3(pf, pt));
    float3 cand[kTeacherCandidates];
    cand[0] = pf;
    cand[1] = pt;
    cand[2] = middle;
    cand[3] = middle;
    double m[3][4] = {};
    for (int side = 0; side < 2; ++side) {
        const uint32_t v = side ? to : from;
        for (uint32_t k = in.face_offsets[v]; k < in.face_offsets[v + 1]; ++k) {
            const uint32_t f = in.faces[k];
            const float3 fa = xyz(in.positions[in.indices[3 * f]]), fb = xyz(in.positions[in.indices[3 * f + 1]]),
                         fc = xyz(in.positions[in.indices[3 * f + 2]]);
            const float3 cr = cross3(sub3(fb, fa), sub3(fc, fa));
            const float3 n = norm3(cr);
            const double length = len3(cr), w[3] = {n.x, n.y, n.z}, d = dot3(n, sub3(fa, middle));
            for (int r = 0; r < 3; ++r) {
                for (int c = 0; c < 3; ++c) m[r][c] += w[r] * w[c] * length;
                m[r][3] += w[r] * d * length;
            }
        }
    }
    bool solved = true;
    for (int col = 0; col < 3; ++col) {
        int pivot = col;
        for (int r = col + 1; r < 3; ++r)
            if (fabs(m[r][col]) > fabs(m[pivot][col])) pivot = r;
        if (fabs(m[pivot][col]) < 1e-15) {
            solved = false;
            break;
        }
        for (int j = col; j < 4; ++j) {
            const double t = m[col][j];
            m[col][j] = m[pivot][j];
            m[pivot][j] = t;
        }
        const double d = m[col][col];
        for (int j = col; j < 4; ++j) m[col][j] /= d;
        for (int r = 0; r < 3; ++r)
            if (r != col) {
                const double factor = m[r][col];
                for (int j = col; j < 4; ++j) m[r][j] -= factor * m[col][j];
            }
    }
    if (solved)
        cand[3] = add3(middle, make_float3(float(fmin(scale, fmax(-scale, m[0][3]))),
                                           float(fmin(scale, fmax(-scale, m[1][3]))),
                                           float(fmin(scale, fmax(-scale, m[2][3])))));
    for (int k = 4; k < 10; ++k) {
        float delta[3] = {0.f, 0.f, 0.f};
        delta[(k - 4) / 2] = float(((k - 4) % 2 ? -.25 : .25) * scale);
        cand[k] = add3(middle, make_float3(delta[0], delta[1], delta[2]));
    }
    for (int k = 0; k < kTeacherCandidates; ++k) {
        const size_t o = (size_t(i) * kTeacherCandidates + k) * 3;
        in.positions_out[o] = cand[k].x;
        in.positions_out[o + 1] = cand[k].y;
        in.positions_out[o + 2] = cand[k].z;
        const float3 off = mul3(sub3(cand[k], middle), 1 / scale);
        in.targets_out[o] = clamp1(off.x);
        in.targets_out[o + 1] = clamp1(off.y);
        in.targets_out[o + 2] = clamp1(off.z);
    }
}

}  // namespace

void launch_features(const FeatureInput& in, hipStream_t stream) {
    if (!in.vertices) return;
    features_kernel<<<dim3((in.vertices + 255) / 256), dim3(256), 0, stream>>>(in);
}

void launch_forward_fp32(const float* x, GraphView g, const float* weights, float* h_a, float* h_b, float* z,
                         float* out, hipStream_t stream) {
    const float* h = x;
    uint32_t width = kFeatures;
    float* targets[3] = {h_a, h_b, h_a};
    for (int l = 0; l < 3; ++l) {
        const size_t n = size_t(g.vertices) * width / 4;
        if (n) aggregate_fp32<<<dim3(uint32_t((n + 255) / 256)), dim3(256), 0, stream>>>(h, g, width, z);
        if (g.vertices)
            compensated_linear<8><<<comp_grid(g.vertices, kHidden, 8), dim3(256), 0, stream>>>(
                z, nullptr, nullptr, weights + layer_offset(l), weights + layer_offset(l) + kHidden * kLayerIn[l],
                targets[l], g.vertices, kLayerIn[l], kHidden, true);
        h = targets[l];
        width = kHidden;
    }
    if (g.actions)
        compensated_linear<8><<<comp_grid(g.actions, kHidden, 8), dim3(256), 0, stream>>>(
            h, g.src, g.neighbors, weights + layer_offset(3), weights + layer_offset(3) + kHidden * kEdgeIn, z,
            g.actions, kEdgeIn, kHidden, true);
    if (g.actions)
        compensated_linear<4><<<comp_grid(g.actions, kOutputs, 4), dim3(256), 0, stream>>>(
            z, nullptr, nullptr, weights + layer_offset(4), weights + layer_offset(4) + kOutputs * kHidden, out,
            g.actions, kHidden, kOutputs, false);
}

// LDS of gemm_bf16<K>: the 128 x K weight block with row padding.
size_t gemm_bf16_lds(uint32_t k) { return size_t(kHidden) * (k + kPad) * 2; }

void launch_convert_bf16(const float* x, __bf16* y, size_t n, hipStream_t stream) {
    if (!n) return;
    to_bf16<<<dim3(uint32_t(std::min<size_t>((n + 255) / 256, 4096))), dim3(256), 0, stream>>>(x, y, n);
}

void launch_offset_copy(const uint32_t* x, uint32_t* y, size_t n, uint32_t add, hipStream_t stream) {
    if (!n) return;
    offset_copy<<<dim3(uint32_t(std::min<size_t>((n + 255) / 256, 4096))), dim3(256), 0, stream>>>(x, y, n, add);
}

void launch_aggregate_fp32(const float* h, GraphView g, uint32_t width, float* z, hipStream_t stream) {
    if (width % 4) throw std::invalid_argument("aggregate_fp32: width must be a multiple of 4");
    const size_t n = size_t(g.vertices) * width / 4;
    if (n) aggregate_fp32<<<dim3(uint32_t((n + 255) / 256)), dim3(256), 0, stream>>>(h, g, width, z);
}

void launch_compensated(const float* input, const uint32_t* src, const uint32_t* dst, const float* weights,
                        const float* bias, float* output, uint32_t rows, uint32_t in, uint32_t out, bool relu,
                        hipStream_t stream) {
    if (in > kCompMaxIn) throw std::invalid_argument("compensated_linear: in must be at most 256");
    const int per = out % 8 == 0 ? 8 : out % 4 == 0 ? 4 : 1;
    if (!rows || !out) return;
    const dim3 grid = comp_grid(rows, out, per);
    if (per == 8)
        compensated_linear<8><<<grid, dim3(256), 0, stream>>>(input, src, dst, weights, bias, output, rows, in, out, relu);
    else if (per == 4)
        compensated_linear<4><<<grid, dim3(256), 0, stream>>>(input, src, dst, weights, bias, output, rows, in, out, relu);
    else
        compensated_linear<1><<<grid, dim3(256), 0, stream>>>(input, src, dst, weights, bias, output, rows, in, out, relu);
}

static void configure_bf16() {
    static bool configured = false;
    if (configured) return;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&gemm_bf16<2 * kFeatures>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(gemm_bf16_lds(2 * kFeatures)));
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&gemm_bf16<2 * kHidden>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(gemm_bf16_lds(2 * kHidden)));
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&gemm_bf16<kHidden>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(gemm_bf16_lds(kHidden)));
    configured = true;
}

// Two persistent 256-thread blocks per CU.
constexpr uint32_t kGemmBlocks = 512;

void launch_mp_bf16(int layer, const __bf16* h, GraphView g, const __bf16* wq, const float* weights, __bf16* out,
                    __bf16* z, hipStream_t stream) {
    configure_bf16();
    if (!g.vertices) return;
    if (!h || !wq || !out || !z) throw std::invalid_argument("null buffer for a BF16 layer");
    const uint32_t tiles = (g.vertices + kTileRows - 1) / kTileRows;
    const dim3 grid(std::min(kGemmBlocks, tiles), 1);
    const float* bias = weights + layer_offset(layer) + kHidden * kLayerIn[layer];
    if (layer == 0) {
        const size_t n = size_t(g.vertices) * (kFeatures / 8);
        aggregate_bf16<kFeatures><<<dim3(uint32_t((n + 255) / 256)), dim3(256), 0, stream>>>(h, g, z);
        gemm_bf16<2 * kFeatures><<<grid, dim3(256), gemm_bf16_lds(2 * kFeatures), stream>>>(
            z, g.vertices, wq + bf16_offset(0), 2 * kFeatures, 0, bias, true, out, kHidden);
    } else {
        const size_t n = size_t(g.vertices) * (kHidden / 8);
        aggregate_bf16<kHidden><<<dim3(uint32_t((n + 255) / 256)), dim3(256), 0, stream>>>(h, g, z);
        gemm_bf16<2 * kHidden><<<grid, dim3(256), gemm_bf16_lds(2 * kHidden), stream>>>(
            z, g.vertices, wq + bf16_offset(layer), 2 * kHidden, 0, bias, true, out, kHidden);
    }
}

void launch_edge_bf16(const __bf16* h, GraphView g, const __bf16* wq, const float* weights, float* out,
                      float* hidden_out, __bf16* p, hipStream_t stream) {
    configure_bf16();
    if (!g.actions || !g.vertices) return;
    if (!h || !wq || !out || !p) throw std::invalid_argument("null buffer for the BF16 edge layer");
    // p = h [W_e(:, :128) | W_e(:, 128:)]^T: two 128-column blocks of W_e.
    const uint32_t tiles = (g.vertices + kTileRows - 1) / kTileRows;
    gemm_bf16<kHidden><<<dim3(std::min(kGemmBlocks, tiles), 2), dim3(256), gemm_bf16_lds(kHidden), stream>>>(
        h, g.vertices, wq + bf16_offset(3), kEdgeIn, kHidden, nullptr, false, p, 2 * kHidden);
    const size_t threads = size_t(g.actions) * 16;
    edge_bf16<<<dim3(uint32_t((threads + 255) / 256)), dim3(256), 0, stream>>>(
        p, g, weights + layer_offset(3) + kHidden * kEdgeIn, weights + layer_offset(4),
        weights + layer_offset(4) + kOutputs * kHidden, out, hidden_out);
}

void launch_forward_bf16(const __bf16* x, GraphView g, const __bf16* wq, const float* weights, __bf16* h_a,
                         __bf16* h_b, __bf16* z, __bf16* p, float* out, hipStream_t stream) {
    launch_mp_bf16(0, x, g, wq, weights, h_a, z, stream);
    launch_mp_bf16(1, h_a, g, wq, weights, h_b, z, stream);
    launch_mp_bf16(2, h_b, g, wq, weights, h_a, z, stream);
    launch_edge_bf16(h_a, g, wq, weights, out, nullptr, p, stream);
}

void launch_scores(const ScoreInput& in, hipStream_t stream) {
    if (!in.actions) return;
    score_kernel<<<dim3((in.actions + 255) / 256), dim3(256), 0, stream>>>(in);
}

size_t sort_workspace(uint32_t n) {
    size_t bytes = 0;
    (void)radix_sort_pairs_descending(nullptr, bytes, nullptr, nullptr, nullptr, nullptr, n, nullptr);
    return bytes;
}

hipError_t launch_sort(void* temp, size_t temp_bytes, const float* keys_in, float* keys_out, const uint32_t* order_in,
                       uint32_t* order_out, uint32_t n, hipStream_t stream) {
    return radix_sort_pairs_descending(temp, temp_bytes, keys_in, keys_out, order_in, order_out, n, stream);
}

void launch_teacher(const TeacherInput& in, hipStream_t stream) {
    if (!in.count) return;
    teacher_kernel<<<dim3((in.count + 127) / 128), dim3(128), 0, stream>>>(in);
}

void launch_decode(const DecodeInput& in, hipStream_t stream) {
    if (!in.count) return;
    decode_kernel<<<dim3((in.count + 255) / 256), dim3(256), 0, stream>>>(in);
}

}  // namespace blitz_hip
