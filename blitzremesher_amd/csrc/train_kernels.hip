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

#include <string>

#include <stdexcept>

#include "train_kernels.hpp"
#include "device_sort.hpp"

namespace blitz_hip {
namespace {

// Zeroes a[0..na) and b[0..nb).
__global__ void fill_u32(uint32_t* __restrict__ a, uint32_t na, uint32_t* __restrict__ b, uint32_t nb) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < na + nb; i += gridDim.x * blockDim.x) {
        if (i < na)
            a[i] = 0;
        else
            b[i - na] = 0;
    }
}

void hip_check_scan(hipError_t e) {
    if (e != hipSuccess) throw std::runtime_error(std::string("gradient scan: ") + hipGetErrorString(e));
}

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
    // One partial per wave, summed in a fixed order by reduce_parts.
    if ((threadIdx.x & 63) == 0) loss[blockIdx.x * 4 + (threadIdx.x >> 6)] = l;
}

// out = the sum of parts[0..n) in a fixed order (thread t takes t, t + 256,
// ... in turn, then a fixed tree), so the result does not depend on launch
// timing.
__global__ __launch_bounds__(256) void reduce_parts(const float* __restrict__ parts, uint32_t n, float* out) {
    __shared__ float sum[256];
    float s = 0.f;
    for (uint32_t i = threadIdx.x; i < n; i += 256) s += parts[i];
    sum[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sum[threadIdx.x] += sum[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sum[0];
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
// at the write. The waves' tiles are combined in LDS in wave order and the
// block writes its slice's tile to `partial`; gemm_at_b_reduce adds the
// slices into C in slice order, so the result does not depend on which
// block finishes first. With `sums`, the column sums of (A masked) (the bias
// gradients) are accumulated from the same loaded values and written per
// slice by the blocks of the first column tile of C.
template <bool kMasked>
__global__ __launch_bounds__(256) void gemm_at_b(const float* __restrict__ a, uint32_t lda,
                                                 const float* __restrict__ a_mask, const float* __restrict__ b,
                                                 uint32_t ldb, float* __restrict__ partial, bool with_sums,
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
    // The block's share: its whole 64 x 64 tile (local row mi - tm, column
    // nj - tn) in its own slot of `partial`, then, for the first column tile
    // of C, the 64 column sums of A after all tiles. gemm_at_b_reduce adds
    // the slices in slice order, so the result is the same bits every run.
    const uint32_t m_tiles = (m + 63) / 64;
    float* mine = partial + (size_t(slice) * tiles + tile) * 4096;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            const floatx4 v = acc[i][j] + part[0][i * 4 + j][lane] + part[1][i * 4 + j][lane] +
                              part[2][i * 4 + j][lane];
            const uint32_t lj = 4 * (lane % 16) + j;
            for (int q = 0; q < 4; ++q) mine[(4 * (4 * (lane / 16) + q) + i) * 64 + lj] = v[q];
        }
    if (with_sums && tn == 0 && lane < 16) {
        // After the tiles of every slice (gridDim.x = tiles x slices blocks).
        float* col_sums = partial + size_t(gridDim.x) * 4096 + (size_t(slice) * m_tiles + tm / 64) * 64;
        for (int i = 0; i < 4; ++i)
            col_sums[4 * lane + i] = col[i] + part_sums[0][4 * lane + i] + part_sums[1][4 * lane + i] +
                                     part_sums[2][4 * lane + i];
    }
}

// c[mi][nj] += the slices' partial tiles in slice order; sums[mi] += the
// slices' column sums in slice order (when sums is set).
__global__ __launch_bounds__(256) void gemm_at_b_reduce(const float* __restrict__ partial, float* __restrict__ c,
                                                        float* __restrict__ sums, uint32_t m, uint32_t n,
                                                        uint32_t slices) {
    const uint32_t n_tiles = (n + 63) / 64, tiles = ((m + 63) / 64) * n_tiles, m_tiles = (m + 63) / 64;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < size_t(m) * n) {
        const uint32_t mi = uint32_t(i / n), nj = uint32_t(i % n);
        const uint32_t tile = (mi / 64) * n_tiles + nj / 64, local = (mi % 64) * 64 + nj % 64;
        float v = 0.f;
        for (uint32_t sl = 0; sl < slices; ++sl) v += partial[(size_t(sl) * tiles + tile) * 4096 + local];
        c[i] += v;
    }
    if (sums && i < m) {
        const float* col_sums = partial + size_t(tiles) * 4096 * slices;
        float v = 0.f;
        for (uint32_t sl = 0; sl < slices; ++sl) v += col_sums[(size_t(sl) * m_tiles + i / 64) * 64 + i % 64];
        sums[i] += v;
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

// The transpose of gather_edge without float atomics: entry e = 2 r + half
// of the batch reads h(src[r]) (half 0) or h(dst[r]) (half 1); every vertex
// gathers the gradients of its entries in ascending e. edge_count and
// edge_fill bucket the entries by vertex (integer atomics; the order inside
// a bucket is then put right by edge_gather).
__global__ void edge_count(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint32_t count,
                           uint32_t* __restrict__ counts) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 2 * count) return;
    atomicAdd(&counts[(e & 1) ? dst[e >> 1] : src[e >> 1]], 1u);
}

__global__ void edge_fill(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint32_t count,
                          const uint32_t* __restrict__ offsets, uint32_t* __restrict__ cursor,
                          uint32_t* __restrict__ list) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 2 * count) return;
    const uint32_t v = (e & 1) ? dst[e >> 1] : src[e >> 1];
    list[offsets[v] + atomicAdd(&cursor[v], 1u)] = e;
}

// One wave per vertex: dh[v][c] = sum over its entries e in ascending order
// of dz[e / 2][(e % 2) * 128 + c]; vertices without entries get zeros. Lane l
// owns channels l and l + 64.
__global__ __launch_bounds__(256) void edge_gather(const float* __restrict__ dz, const uint32_t* __restrict__ offsets,
                                                   const uint32_t* __restrict__ list, uint32_t vertices,
                                                   float* __restrict__ dh) {
    const uint32_t lane = threadIdx.x & 63, v = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (v >= vertices) return;
    const uint32_t b = offsets[v], len = offsets[v + 1] - b;
    float a0 = 0.f, a1 = 0.f;
    const auto take = [&](uint32_t e) {
        const float* row = dz + size_t(e >> 1) * kEdgeIn + (e & 1) * kHidden;
        a0 += row[lane];
        a1 += row[lane + 64];
    };
    if (len <= 64) {
        // Rank of each entry among the bucket's (all distinct), then the
        // entries in rank order.
        const uint32_t code = lane < len ? list[b + lane] : 0xFFFFFFFFu;
        uint32_t rank = 0;
        for (uint32_t k = 0; k < len; ++k) rank += uint32_t(__shfl(int(code), int(k))) < code;
        for (uint32_t k = 0; k < len; ++k) {
            const uint64_t owner = __ballot(lane < len && rank == k);
            take(uint32_t(__shfl(int(code), __ffsll(static_cast<unsigned long long>(owner)) - 1)));
        }
    } else {
        // Long buckets: repeatedly the smallest entry above the last one.
        uint32_t last = 0;
        for (uint32_t k = 0; k < len; ++k) {
            uint32_t best = 0xFFFFFFFFu;
            for (uint32_t j = lane; j < len; j += 64) {
                const uint32_t e = list[b + j];
                if ((k == 0 || e > last) && e < best) best = e;
            }
            for (int o = 32; o > 0; o >>= 1) best = min(best, uint32_t(__shfl_xor(int(best), o)));
            take(best);
            last = best;
        }
    }
    dh[size_t(v) * kHidden + lane] = a0;
    dh[size_t(v) * kHidden + lane + 64] = a1;
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
    if ((threadIdx.x & 63) == 0) sum[blockIdx.x * 4 + (threadIdx.x >> 6)] = s;  // one partial per wave
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

uint32_t loss_parts(const LossInput& in) { return (in.count + in.pool - 1) / in.pool * 4; }

void launch_loss(const float* y, const uint8_t* labels, const float* targets, float* dy, float* parts, float* loss,
                 const LossInput& in, hipStream_t stream) {
    if (!in.count) return;
    loss_kernel<<<dim3((in.count + in.pool - 1) / in.pool), dim3(256), 0, stream>>>(y, labels, targets, dy, parts, in);
    reduce_parts<<<1, 256, 0, stream>>>(parts, loss_parts(in), loss);
}

namespace {
// About 512 blocks (two per CU) of at least 256 rows each; the slice count is
// a multiple of 8 (one share per XCD).
uint32_t gemm_at_b_slices(uint32_t rows, uint32_t m, uint32_t n) {
    const uint32_t tiles = ((m + 63) / 64) * ((n + 63) / 64);
    const uint32_t wanted = std::max<uint32_t>(1, std::min<uint32_t>((rows + 255) / 256, 512 / tiles));
    return (wanted + 7) / 8 * 8;
}
}  // namespace

size_t gemm_at_b_workspace(uint32_t rows, uint32_t m, uint32_t n) {
    const uint32_t tiles = ((m + 63) / 64) * ((n + 63) / 64), slices = gemm_at_b_slices(rows, m, n);
    return size_t(slices) * tiles * 4096 + size_t(slices) * ((m + 63) / 64) * 64;
}

void launch_gemm_at_b(const float* a, uint32_t lda, const float* a_mask, const float* b, uint32_t ldb, float* c,
                      float* sums, float* work, uint32_t rows, uint32_t m, uint32_t n, hipStream_t stream) {
    if (!rows || !m || !n) return;
    if ((m | n | lda | ldb) % 4) throw std::invalid_argument("gemm_at_b: m, n and row strides must be multiples of 4");
    const uint32_t tiles = ((m + 63) / 64) * ((n + 63) / 64);
    const uint32_t slices = gemm_at_b_slices(rows, m, n);
    const uint32_t per = (rows + slices - 1) / slices;
    const dim3 grid(tiles * slices);
    if (a_mask)
        gemm_at_b<true><<<grid, dim3(256), 0, stream>>>(a, lda, a_mask, b, ldb, work, sums != nullptr, rows, m, n, per);
    else
        gemm_at_b<false><<<grid, dim3(256), 0, stream>>>(a, lda, a_mask, b, ldb, work, sums != nullptr, rows, m, n, per);
    const size_t elements = std::max<size_t>(size_t(m) * n, m);
    gemm_at_b_reduce<<<dim3(uint32_t((elements + 255) / 256)), dim3(256), 0, stream>>>(work, c, sums, m, n, slices);
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

void launch_gather_edge_backward(const float* dz, const uint32_t* src, const uint32_t* dst, uint32_t count,
                                 uint32_t vertices, uint32_t* counts, uint32_t* offsets, uint32_t* cursor,
                                 uint32_t* list, void* scan_temp, size_t scan_bytes, float* dh, hipStream_t stream) {
    fill_u32<<<dim3(std::min<uint32_t>(1024, (2 * vertices + 2 + 255) / 256)), dim3(256), 0, stream>>>(
        counts, vertices + 1, cursor, vertices);
    const uint32_t entries = 2 * count;
    if (entries) {
        edge_count<<<dim3((entries + 255) / 256), dim3(256), 0, stream>>>(src, dst, count, counts);
        hip_check_scan(exclusive_sum(scan_temp, scan_bytes, counts, offsets, vertices + 1, stream));
        edge_fill<<<dim3((entries + 255) / 256), dim3(256), 0, stream>>>(src, dst, count, offsets, cursor, list);
    } else {
        fill_u32<<<dim3(std::min<uint32_t>(1024, (vertices + 1 + 255) / 256)), dim3(256), 0, stream>>>(
            offsets, vertices + 1, cursor, 0);
    }
    edge_gather<<<dim3((vertices + 3) / 4), dim3(256), 0, stream>>>(dz, offsets, list, vertices, dh);
}

size_t gather_edge_scan_bytes(uint32_t vertices) {
    size_t bytes = 0;
    (void)exclusive_sum(nullptr, bytes, static_cast<const uint32_t*>(nullptr), nullptr, vertices + 1, nullptr);
    return std::max<size_t>(bytes, 16);
}

void launch_aggregate_backward(const float* t, GraphView g, uint32_t width, float* dh, hipStream_t stream) {
    if (width % 4) throw std::invalid_argument("aggregate_backward: width must be a multiple of 4");
    const size_t n = size_t(g.vertices) * width / 4;
    if (!n) return;
    aggregate_backward<<<dim3(uint32_t((n + 255) / 256)), dim3(256), 0, stream>>>(t, g, width, dh);
}

void launch_norm(const float* g, uint32_t n, float* parts, float* sum, hipStream_t stream) {
    const uint32_t blocks = std::min<uint32_t>((n + 255) / 256, 1024);
    norm_kernel<<<dim3(blocks), dim3(256), 0, stream>>>(g, n, parts);
    reduce_parts<<<1, 256, 0, stream>>>(parts, blocks * 4, sum);
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
