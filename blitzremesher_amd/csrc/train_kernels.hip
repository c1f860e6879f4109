// Training kernels of the edge-graph scorer: loss, backward, optimizer.
//
// Loss over a batch of labelled actions grouped into pools of P consecutive
// entries (the reference's per-state pools; the joint-scope terms of
// training/update_cuda.cu:256-330 applied to this network's four outputs
// [logit, placement x3]): pairwise hinge on the logit between preferred and
// other valid actions of a pool (margin m, normalised by the number of pairs
// over the batch), an L2 penalty of weight penalty/valid/3 on the logit of
// valid actions, and a smooth-L1 (delta 1) loss, normalised by 3 x the number
// of placed entries, on the three placement outputs of placed entries.
// Gradient clipping by global norm and AdamW restate
// training/update_cuda.cu:338-393 operation for operation.
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "train_kernels.hpp"

namespace blitz_hip {
namespace {

// Label bits of one batch entry.
constexpr uint8_t kValid = 1, kPreferred = 2, kPlaced = 4;

__global__ __launch_bounds__(256) void loss_kernel(const float* __restrict__ y, const uint8_t* __restrict__ labels,
                                                   const float* __restrict__ targets, float* __restrict__ dy,
                                                   float* __restrict__ loss, LossInput in) {
    const uint32_t pool0 = blockIdx.x * in.pool;
    const uint32_t j = pool0 + threadIdx.x;
    float l = 0.f;
    if (threadIdx.x < in.pool && j < in.count) {
        const uint8_t label = labels[j];
        float g[4] = {0, 0, 0, 0};
        const float* p = y + size_t(j) * 4;
        if (label & kValid) {
            const bool preferred = label & kPreferred;
            const float score = p[0];
            const float inv = 1.f / float(in.pairs > 0 ? in.pairs : 1u);
            const uint32_t end = min(pool0 + in.pool, in.count);
            for (uint32_t k = pool0; k < end; ++k) {
                const uint8_t other = labels[k];
                if ((other & kValid) && bool(other & kPreferred) != preferred) {
                    const float v = in.margin + (preferred ? y[size_t(k) * 4] - score : score - y[size_t(k) * 4]);
                    if (v > 0) {
                        g[0] += (preferred ? -1.f : 1.f) * inv;
                        if (preferred) l += v * inv;
                    }
                }
            }
            const float w = in.penalty * (1.f / float(in.valid > 0 ? in.valid : 1u)) / 3;
            l += w * score * score;
            g[0] += 2 * w * score;
        }
        if (label & kPlaced) {
            const float inv = in.placement_weight / (float(in.placed > 0 ? in.placed : 1u) * 3.f);
            for (int c = 0; c < 3; ++c) {
                const float delta = p[1 + c] - targets[size_t(j) * 3 + c];
                const float a = fabsf(delta);
                l += (a < 1 ? .5f * delta * delta : a - .5f) * inv;
                g[1 + c] = fminf(1.f, fmaxf(-1.f, delta)) * inv;
            }
        }
        for (int c = 0; c < 4; ++c) dy[size_t(j) * 4 + c] = g[c];
    }
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
    if ((threadIdx.x & 63) == 0 && l != 0.f) atomicAdd(loss, l);
}

// v_mfma_f32_16x16x4f32 (FP32 products and sums) operand layout, wave64:
// lane l supplies A(i = l%16, k = l/16) and B(k = l/16, j = l%16); element q
// of lane l's accumulator is D(i = 4*(l/16) + q, j = l%16).
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ floatx4 mfma4(float a, float b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float4 ld4(const float* p) { return *reinterpret_cast<const float4*>(p); }

__device__ __forceinline__ float4 add4(float4 a, float4 b) {
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

__device__ __forceinline__ float part_of(const float4& v, int s) {
    return s == 0 ? v.x : s == 1 ? v.y : s == 2 ? v.z : v.w;
}

// An operand value: zero out of range and, masked, where mask <= 0 (the
// ReLU derivative). Loads are issued unconditionally (out-of-range lanes read
// element 0) one step before this select, so no load waits behind a branch.
__device__ __forceinline__ float keep(bool in, float v, float mask) { return in && mask > 0.f ? v : 0.f; }

// C[M][N] += sum_r A(r, m) * B(r, n) on the matrix cores. A and B are
// row-major with row strides lda/ldb; kMasked zeroes A(r, m) where
// a_mask(r, m) <= 0, read with A's layout. A block owns a 64 x 64 tile of C
// and a slice of rows, a quarter per wave; four rows make one MFMA k-step and
// sixteen rows one iteration, whose loads are issued an iteration ahead.
// Lane l loads four consecutive columns 4*(l%16) .. +3 of row l/16 of A and
// of B and feeds component s to sub-tile s, so sub-tile s holds tile columns
// 4*i + s (rows of C) and 4*j + s (columns of C): a fixed permutation undone
// at the write. The waves' tiles are combined in LDS and added atomically.
// With `sums`, the column sums of (A masked) (the bias gradients) are
// accumulated from the same loaded values and added by the blocks of the
// first column tile of C.
template <bool kMasked>
__global__ __launch_bounds__(256) void gemm_at_b(const float* __restrict__ a, uint32_t lda,
                                                 const float* __restrict__ a_mask, const float* __restrict__ b,
                                                 uint32_t ldb, float* __restrict__ c, float* __restrict__ sums,
                                                 uint32_t rows, uint32_t m, uint32_t n, uint32_t rows_per_slice) {
    __shared__ floatx4 part[3][16][64];
    __shared__ float part_sums[3][64];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // Blocks are dealt to the eight XCDs round-robin (block b on XCD b % 8);
    // all tiles of one row slice go to the same XCD, one after another, so
    // their waves walk the same rows together through that XCD's L2.
    const uint32_t n_tiles = (n + 63) / 64, tiles = ((m + 63) / 64) * n_tiles;
    const uint32_t local = blockIdx.x / 8, tile = local % tiles;
    const uint32_t slice = (local / tiles) * 8 + blockIdx.x % 8;
    const uint32_t tm = tile / n_tiles * 64, tn = tile % n_tiles * 64;
    const uint32_t s0 = min(rows, slice * rows_per_slice), s1 = min(rows, s0 + rows_per_slice);
    const uint32_t quarter = ((rows_per_slice + 63) / 64) * 16;  // rows per wave, a multiple of 16
    const uint32_t r0 = min(s1, s0 + wave * quarter), r1 = min(s1, r0 + quarter);
    const uint32_t ma = tm + 4 * (lane % 16), nb = tn + 4 * (lane % 16);
    const bool a_in = ma < m, b_in = nb < n;
    struct Rows {
        float4 a[4], mask[4], b[4];
    };
    auto fetch = [&](uint32_t base, Rows& x) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const uint32_t row = base + 4 * s + lane / 16;
            const bool in = row < r1;
            const size_t at = in && a_in ? size_t(row) * lda + ma : 0;
            x.a[s] = ld4(a + at);
            if (kMasked) x.mask[s] = ld4(a_mask + at);
            x.b[s] = ld4(b + (in && b_in ? size_t(row) * ldb + nb : 0));
        }
    };
    floatx4 acc[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    float col[4] = {0.f, 0.f, 0.f, 0.f};  // column sums of (A masked): bias gradients
    Rows cur, next;
    fetch(r0, cur);
    for (uint32_t r = r0; r < r1; r += 16) {
        fetch(r + 16, next);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const uint32_t row = r + 4 * s + lane / 16;
            const bool ia = row < r1 && a_in, ib = row < r1 && b_in;
            float av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                av[i] = keep(ia, part_of(cur.a[s], i), kMasked ? part_of(cur.mask[s], i) : 1.f);
                bv[i] = ib ? part_of(cur.b[s], i) : 0.f;
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                col[i] += av[i];
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = mfma4(av[i], bv[j], acc[i][j]);
            }
        }
        cur = next;
    }
    // Lanes l, l+16, l+32, l+48 hold the same four columns of A.
    for (int i = 0; i < 4; ++i) {
        col[i] += __shfl_xor(col[i], 16);
        col[i] += __shfl_xor(col[i], 32);
    }
    if (wave) {
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) part[wave - 1][i * 4 + j][lane] = acc[i][j];
        if (lane < 16)
            for (int i = 0; i < 4; ++i) part_sums[wave - 1][4 * lane + i] = col[i];
    }
    __syncthreads();
    if (wave) return;
    // The column sums of A are added once, by the first column tile of C.
    if (sums && tn == 0 && lane < 16 && a_in)
        for (int i = 0; i < 4; ++i) {
            const float v = col[i] + part_sums[0][4 * lane + i] + part_sums[1][4 * lane + i] +
                            part_sums[2][4 * lane + i];
            if (v != 0.f) atomicAdd(&sums[ma + i], v);
        }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            const floatx4 v = acc[i][j] + part[0][i * 4 + j][lane] + part[1][i * 4 + j][lane] +
                              part[2][i * 4 + j][lane];
            const uint32_t nj = tn + 4 * (lane % 16) + j;
            for (int q = 0; q < 4; ++q) {
                const uint32_t mi = tm + 4 * (4 * (lane / 16) + q) + i;
                if (mi < m && nj < n && v[q] != 0.f) atomicAdd(&c[size_t(mi) * n + nj], v[q]);
            }
        }
}

// T[R][N] = (A masked)[R][K] x W[K][N] (W row-major with K rows of N) on the
// matrix cores. A wave owns 64 rows x 64 columns of T; a block's waves take
// consecutive column tiles of the same 64 rows, so they share the reads of A
// through the CU's caches. Per 16 columns of K,
// lane l loads four consecutive A columns 4*(l/16) .. +3 of row l%16 of each
// 16-row sub-tile and feeds component s to k-step s, so k-step s covers K
// columns 4*g + s (g = l/16) — the same permutation applied to W's rows. The
// next 16 columns' loads are issued before this block's 64 MFMAs.
template <bool kMasked>
__global__ __launch_bounds__(256) void gemm_a_w(const float* __restrict__ a, const float* __restrict__ a_mask,
                                                const float* __restrict__ w, float* __restrict__ t, uint32_t rows,
                                                uint32_t k_dim, uint32_t n) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t tr = blockIdx.x * 64, tn = (blockIdx.y * (blockDim.x / 64) + threadIdx.x / 64) * 64;
    if (tn >= n) return;
    const uint32_t g = lane / 16, col = lane % 16;
    struct Block {
        float4 a[4], mask[4];
        float w[4][4];
    };
    auto fetch = [&](uint32_t k0, Block& x) {
        const uint32_t ka = k0 + 4 * g;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t row = tr + 16 * i + col;
            const size_t at = row < rows && ka < k_dim ? size_t(row) * k_dim + ka : 0;
            x.a[i] = ld4(a + at);
            if (kMasked) x.mask[i] = ld4(a_mask + at);
        }
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t nj = tn + 16 * j + col;
                x.w[s][j] = w[ka < k_dim && nj < n ? size_t(ka + s) * n + nj : 0];
            }
    };
    floatx4 acc[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    Block cur, next;
    fetch(0, cur);
    for (uint32_t k0 = 0; k0 < k_dim; k0 += 16) {
        fetch(k0 + 16, next);
        const uint32_t ka = k0 + 4 * g;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            float wv[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) wv[j] = ka < k_dim && tn + 16 * j + col < n ? cur.w[s][j] : 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const bool in = tr + 16 * i + col < rows && ka < k_dim;
                const float x = keep(in, part_of(cur.a[i], s), kMasked ? part_of(cur.mask[i], s) : 1.f);
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = mfma4(x, wv[j], acc[i][j]);
            }
        }
        cur = next;
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            const uint32_t nj = tn + 16 * j + col;
            for (int q = 0; q < 4; ++q) {
                const uint32_t ri = tr + 16 * i + 4 * g + q;
                if (ri < rows && nj < n) t[size_t(ri) * n + nj] = acc[i][j][q];
            }
        }
}

// Endpoints of the batch's actions.
__global__ void batch_endpoints(const uint32_t* __restrict__ actions, GraphView g, uint32_t* __restrict__ src,
                                uint32_t* __restrict__ dst, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    src[i] = g.src[actions[i]];
    dst[i] = g.neighbors[actions[i]];
}

// z[r] = [h(src[r]) ; h(dst[r])], the edge layer's input.
__global__ void gather_edge(const float* __restrict__ h, const uint32_t* __restrict__ src,
                            const uint32_t* __restrict__ dst, float* __restrict__ z, uint32_t count) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(count) * kEdgeIn) return;
    const uint32_t r = uint32_t(i / kEdgeIn), c = uint32_t(i % kEdgeIn);
    const uint32_t v = c < kHidden ? src[r] : dst[r];
    z[i] = h[size_t(v) * kHidden + (c % kHidden)];
}

// dz of the edge input [h(src) ; h(dst)] scattered into dh.
__global__ void scatter_edge(const float* __restrict__ dz, const uint32_t* __restrict__ src,
                             const uint32_t* __restrict__ dst, float* __restrict__ dh, uint32_t count) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(count) * kEdgeIn) return;
    const uint32_t r = uint32_t(i / kEdgeIn), c = uint32_t(i % kEdgeIn);
    const uint32_t v = c < kHidden ? src[r] : dst[r];
    atomicAdd(&dh[size_t(v) * kHidden + (c % kHidden)], dz[i]);
}

// dh[v] = t[v][0:d] + sum over u in N(v) of t[u][d:2d] / deg(u): the
// transpose of the mean aggregation on a symmetric neighbour graph. Four
// channels per thread, neighbours four at a time, adds in CSR order.
__global__ __launch_bounds__(256) void aggregate_backward(const float* __restrict__ t, GraphView g, uint32_t width,
                                                          float* __restrict__ dh) {
    const uint32_t quads = width / 4;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(g.vertices) * quads) return;
    const uint32_t v = uint32_t(i / quads), c = uint32_t(i % quads) * 4;
    auto part = [&](uint32_t u) {
        const uint32_t deg = g.offsets[u + 1] - g.offsets[u];
        const float d = float(deg > 0 ? deg : 1u);
        const float4 x = *reinterpret_cast<const float4*>(t + size_t(u) * 2 * width + width + c);
        return make_float4(x.x / d, x.y / d, x.z / d, x.w / d);
    };
    float4 s = *reinterpret_cast<const float4*>(t + size_t(v) * 2 * width + c);
    const uint32_t e = g.offsets[v + 1];
    uint32_t k = g.offsets[v];
    for (; k + 4 <= e; k += 4) {
        const float4 a0 = part(g.neighbors[k]), a1 = part(g.neighbors[k + 1]), a2 = part(g.neighbors[k + 2]),
                     a3 = part(g.neighbors[k + 3]);
        s = add4(add4(add4(add4(s, a0), a1), a2), a3);
    }
    for (; k < e; ++k) s = add4(s, part(g.neighbors[k]));
    *reinterpret_cast<float4*>(dh + size_t(v) * width + c) = s;
}

__global__ void norm_kernel(const float* __restrict__ g, uint32_t n, float* __restrict__ sum) {
    float s = 0.f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float v = g[i];
        s += isfinite(v) ? v * v : INFINITY;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) atomicAdd(sum, s);
}

// training/update_cuda.cu:378-393 with the clip and bias corrections
// computed on the host from the norm (training/update_cuda.cu:358-376).
__global__ void adamw_kernel(float* __restrict__ value, const float* __restrict__ grad, float* __restrict__ mean,
                             float* __restrict__ variance, uint32_t n, AdamStep s) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = grad[i] * s.clip;
    const float m = mean[i] * .9f + g * .1f, v = variance[i] * .999f + (g * g) * .001f;
    mean[i] = m;
    variance[i] = v;
    value[i] = value[i] * (1 - s.lr * s.decay) - s.correction1 * m / (sqrtf(v) / s.correction2 + 1e-8f);
}

__global__ void bf16_to_fp32(const __bf16* __restrict__ x, float* __restrict__ y, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        y[i] = float(x[i]);
}

}  // namespace

void launch_loss(const float* y, const uint8_t* labels, const float* targets, float* dy, float* loss,
                 const LossInput& in, hipStream_t stream) {
    if (!in.count) return;
    loss_kernel<<<dim3((in.count + in.pool - 1) / in.pool), dim3(256), 0, stream>>>(y, labels, targets, dy, loss, in);
}

void launch_gemm_at_b(const float* a, uint32_t lda, const float* a_mask, const float* b, uint32_t ldb, float* c,
                      float* sums, uint32_t rows, uint32_t m, uint32_t n, hipStream_t stream) {
    if (!rows || !m || !n) return;
    if ((m | n | lda | ldb) % 4) throw std::invalid_argument("gemm_at_b: m, n and row strides must be multiples of 4");
    // About 512 blocks (two per CU) of at least 256 rows each; the slice
    // count is a multiple of 8 (one share per XCD).
    const uint32_t tiles = ((m + 63) / 64) * ((n + 63) / 64);
    const uint32_t wanted = std::max<uint32_t>(1, std::min<uint32_t>((rows + 255) / 256, 512 / tiles));
    const uint32_t slices = (wanted + 7) / 8 * 8;
    const uint32_t per = (rows + slices - 1) / slices;
    const dim3 grid(tiles * slices);
    if (a_mask)
        gemm_at_b<true><<<grid, dim3(256), 0, stream>>>(a, lda, a_mask, b, ldb, c, sums, rows, m, n, per);
    else
        gemm_at_b<false><<<grid, dim3(256), 0, stream>>>(a, lda, a_mask, b, ldb, c, sums, rows, m, n, per);
}

void launch_gemm_a_w(const float* a, const float* a_mask, const float* w, float* t, uint32_t rows, uint32_t k,
                     uint32_t n, hipStream_t stream) {
    if (!rows || !n) return;
    if (k % 4) throw std::invalid_argument("gemm_a_w: k must be a multiple of 4");
    const uint32_t tiles = (n + 63) / 64, waves = std::min(4u, tiles);
    const dim3 grid((rows + 63) / 64, (tiles + waves - 1) / waves), block(64 * waves);
    if (a_mask)
        gemm_a_w<true><<<grid, block, 0, stream>>>(a, a_mask, w, t, rows, k, n);
    else
        gemm_a_w<false><<<grid, block, 0, stream>>>(a, a_mask, w, t, rows, k, n);
}

void launch_batch_endpoints(const uint32_t* actions, GraphView g, uint32_t* src, uint32_t* dst, uint32_t count,
                            hipStream_t stream) {
    if (!count) return;
    batch_endpoints<<<dim3((count + 255) / 256), dim3(256), 0, stream>>>(actions, g, src, dst, count);
}

void launch_gather_edge(const float* h, const uint32_t* src, const uint32_t* dst, float* z, uint32_t count,
                        hipStream_t stream) {
    const size_t n = size_t(count) * kEdgeIn;
    if (!n) return;
    gather_edge<<<dim3(uint32_t((n + 255) / 256)), dim3(256), 0, stream>>>(h, src, dst, z, count);
}

void launch_scatter_edge(const float* dz, const uint32_t* src, const uint32_t* dst, float* dh, uint32_t count,
                         hipStream_t stream) {
    const size_t n = size_t(count) * kEdgeIn;
    if (!n) return;
    scatter_edge<<<dim3(uint32_t((n + 255) / 256)), dim3(256), 0, stream>>>(dz, src, dst, dh, count);
}

void launch_aggregate_backward(const float* t, GraphView g, uint32_t width, float* dh, hipStream_t stream) {
    if (width % 4) throw std::invalid_argument("aggregate_backward: width must be a multiple of 4");
    const size_t n = size_t(g.vertices) * width / 4;
    if (!n) return;
    aggregate_backward<<<dim3(uint32_t((n + 255) / 256)), dim3(256), 0, stream>>>(t, g, width, dh);
}

void launch_norm(const float* g, uint32_t n, float* sum, hipStream_t stream) {
    norm_kernel<<<dim3(std::min<uint32_t>((n + 255) / 256, 1024)), dim3(256), 0, stream>>>(g, n, sum);
}

void launch_adamw(float* value, const float* grad, float* mean, float* variance, uint32_t n, const AdamStep& s,
                  hipStream_t stream) {
    adamw_kernel<<<dim3((n + 255) / 256), dim3(256), 0, stream>>>(value, grad, mean, variance, n, s);
}

void launch_bf16_to_fp32(const __bf16* x, float* y, size_t n, hipStream_t stream) {
    if (!n) return;
    bf16_to_fp32<<<dim3(uint32_t(std::min<size_t>((n + 255) / 256, 4096))), dim3(256), 0, stream>>>(x, y, n);
}

}  // namespace blitz_hip
