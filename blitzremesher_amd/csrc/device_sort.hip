// Stable LSD radix sort, prefix scans and run-length encoding (published
// methods: LSD radix sort by counting, one 8-bit digit per pass; a
// three-phase block scan — block totals, a scan of the totals, block scans
// seeded with the scanned totals).
//
// Radix pass, digit d of every key: (1) each tile of kTile keys counts its
// digits in LDS and writes them digit-major (count[digit * tiles + tile]);
// (2) one exclusive scan of those counts gives every (digit, tile) its first
// output slot; (3) each tile walks its keys again in input order, 256 per
// round, and places each key at its digit's slot plus the number of earlier
// keys of the tile with that digit: within a wavefront from eight ballots
// (the lanes whose digit bits all agree), across the round's four wavefronts
// from their per-digit counts in LDS, across rounds from a running per-digit
// base. Keys therefore keep their input order within a digit, and LSD passes
// from the low digit up give a stable sort.
#include "device_sort.hpp"

#include <algorithm>

namespace blitz_hip {
namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kItems = 8;
constexpr uint32_t kTile = kThreads * kItems;  // keys per tile (radix) and elements per block (scan)
constexpr uint32_t kWaves = kThreads / 64;
constexpr uint32_t kDigits = 256;

size_t align_up(size_t v) { return (v + 255) / 256 * 256; }

template <class K>
__global__ __launch_bounds__(kThreads) void digit_count_kernel(const K* __restrict__ keys, uint32_t n, int shift,
                                                               uint32_t tiles, uint32_t* __restrict__ counts) {
    __shared__ uint32_t hist[kDigits];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t first = blockIdx.x * kTile;
    for (uint32_t r = 0; r < kItems; ++r) {
        const uint32_t i = first + r * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&hist[uint32_t(keys[i] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    counts[size_t(threadIdx.x) * tiles + blockIdx.x] = hist[threadIdx.x];
}

template <class K>
__global__ __launch_bounds__(kThreads) void digit_scatter_kernel(const K* __restrict__ keys_in, K* __restrict__ keys_out,
                                                                 const uint32_t* __restrict__ values_in,
                                                                 uint32_t* __restrict__ values_out, uint32_t n,
                                                                 int shift, uint32_t tiles,
                                                                 const uint32_t* __restrict__ offsets) {
    __shared__ uint32_t base[kDigits];
    __shared__ uint32_t wave_count[kWaves][kDigits];
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    base[t] = offsets[size_t(t) * tiles + blockIdx.x];
    const uint64_t below = (1ull << lane) - 1;
    const uint32_t first = blockIdx.x * kTile;
    for (uint32_t r = 0; r < kItems; ++r) {
        for (uint32_t w = 0; w < kWaves; ++w) wave_count[w][t] = 0;
        __syncthreads();
        const uint32_t i = first + r * kThreads + t;
        const bool live = i < n;
        const K key = live ? keys_in[i] : K(0);
        const uint32_t d = uint32_t(key >> shift) & 0xFFu;
        // Lanes of this wavefront holding the same digit.
        uint64_t peers = __ballot(live);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t ones = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? ones : ~ones;
        }
        const uint32_t rank = uint32_t(__popcll(peers & below));
        if (live && rank == 0) wave_count[wave][d] = uint32_t(__popcll(peers));
        __syncthreads();
        if (live) {
            uint32_t at = base[d] + rank;
            for (uint32_t w = 0; w < wave; ++w) at += wave_count[w][d];
            keys_out[at] = key;
            if (values_out) values_out[at] = values_in ? values_in[i] : i;
        }
        __syncthreads();
        uint32_t add = 0;
        for (uint32_t w = 0; w < kWaves; ++w) add += wave_count[w][t];
        base[t] += add;
        __syncthreads();
    }
}

template <class T>
struct SumOp {
    __device__ __forceinline__ T operator()(T a, T b) const { return a + b; }
};
template <class T>
struct MaxOp {
    __device__ __forceinline__ T operator()(T a, T b) const { return a > b ? a : b; }
};

// Inclusive and exclusive scan of one value per thread over the block (0 is
// the identity of both operators on unsigned values).
template <class T>
struct Scanned {
    T inclusive, exclusive;
};
template <class T, class Op>
__device__ __forceinline__ Scanned<T> block_scan(T v, Op op, T* wave_total) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        const T up = __shfl_up(v, o);
        if (lane >= uint32_t(o)) v = op(v, up);
    }
    const T up = __shfl_up(v, 1);
    const T lane_excl = lane ? up : T(0);
    if (lane == 63) wave_total[wave] = v;
    __syncthreads();
    T before = 0;
    for (uint32_t w = 0; w < wave; ++w) before = op(before, wave_total[w]);
    __syncthreads();
    return {op(before, v), op(before, lane_excl)};
}

template <class T, class Op>
__global__ __launch_bounds__(kThreads) void scan_totals_kernel(const T* __restrict__ in, uint32_t n,
                                                               T* __restrict__ totals, Op op) {
    __shared__ T wave_total[kWaves];
    const uint32_t first = blockIdx.x * kTile + threadIdx.x * kItems;
    T v = 0;
    for (uint32_t k = 0; k < kItems; ++k)
        if (first + k < n) v = op(v, in[first + k]);
    v = block_scan(v, op, wave_total).inclusive;
    if (threadIdx.x == kThreads - 1) totals[blockIdx.x] = v;
}

// Exclusive scan of the block totals in place (one workgroup, any count).
template <class T, class Op>
__global__ __launch_bounds__(kThreads) void scan_carry_kernel(T* totals, uint32_t blocks, Op op) {
    __shared__ T wave_total[kWaves];
    __shared__ T carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t b0 = 0; b0 < blocks; b0 += kThreads) {
        const uint32_t b = b0 + threadIdx.x;
        const T v = b < blocks ? totals[b] : T(0);
        const Scanned<T> s = block_scan(v, op, wave_total);
        const T c = carry;
        __syncthreads();
        if (b < blocks) totals[b] = op(c, s.exclusive);
        if (threadIdx.x == kThreads - 1) carry = op(c, s.inclusive);
        __syncthreads();
    }
}

template <class T, class Op, bool Inclusive>
__global__ __launch_bounds__(kThreads) void scan_apply_kernel(const T* in, T* out, uint32_t n,
                                                              const T* __restrict__ carry, Op op) {
    __shared__ T wave_total[kWaves];
    const uint32_t first = blockIdx.x * kTile + threadIdx.x * kItems;
    T v[kItems];
    T sum = 0;
    for (uint32_t k = 0; k < kItems; ++k) {
        v[k] = first + k < n ? in[first + k] : T(0);
        sum = op(sum, v[k]);
    }
    // Everything before this thread's items: the block's carry and the
    // threads before it.
    T run = op(carry[blockIdx.x], block_scan(sum, op, wave_total).exclusive);
    for (uint32_t k = 0; k < kItems; ++k) {
        if (first + k >= n) break;
        const T next = op(run, v[k]);
        out[first + k] = Inclusive ? next : run;
        run = next;
    }
}

// Short scans (every radix pass of a few tens of thousands of keys scans
// 256 counts per tile): one workgroup walks the array in chunks of 4,096 and
// carries the running total, one launch instead of three.
constexpr uint32_t kSmallThreads = 1024;
constexpr uint32_t kSmallItems = 4;
constexpr uint32_t kSmallScan = 16 * kSmallThreads * kSmallItems;

template <class T, class Op, bool Inclusive>
__global__ __launch_bounds__(kSmallThreads) void scan_small_kernel(const T* in, T* out, uint32_t n, Op op) {
    __shared__ T wave_total[kSmallThreads / 64];
    __shared__ T carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += kSmallThreads * kSmallItems) {
        const uint32_t first = base + threadIdx.x * kSmallItems;
        T v[kSmallItems];
        T sum = 0;
        for (uint32_t k = 0; k < kSmallItems; ++k) {
            v[k] = first + k < n ? in[first + k] : T(0);
            sum = op(sum, v[k]);
        }
        const Scanned<T> s = block_scan(sum, op, wave_total);
        const T c = carry;
        T run = op(c, s.exclusive);
        for (uint32_t k = 0; k < kSmallItems; ++k) {
            if (first + k >= n) break;
            const T next = op(run, v[k]);
            out[first + k] = Inclusive ? next : run;
            run = next;
        }
        __syncthreads();
        if (threadIdx.x == kSmallThreads - 1) carry = op(c, s.inclusive);
        __syncthreads();
    }
}

template <class T, class Op, bool Inclusive>
hipError_t scan(void* temp, size_t& temp_bytes, const T* in, T* out, uint32_t n, hipStream_t stream, Op op) {
    const uint32_t blocks = std::max(1u, (n + kTile - 1) / kTile);
    const size_t need = align_up(size_t(blocks) * sizeof(T));
    if (!temp) {
        temp_bytes = need;
        return hipSuccess;
    }
    if (temp_bytes < need) return hipErrorInvalidValue;
    if (!n) return hipSuccess;
    if (n <= kSmallScan) {
        scan_small_kernel<T, Op, Inclusive><<<1, kSmallThreads, 0, stream>>>(in, out, n, op);
        return hipGetLastError();
    }
    T* totals = static_cast<T*>(temp);
    scan_totals_kernel<T, Op><<<blocks, kThreads, 0, stream>>>(in, n, totals, op);
    scan_carry_kernel<T, Op><<<1, kThreads, 0, stream>>>(totals, blocks, op);
    scan_apply_kernel<T, Op, Inclusive><<<blocks, kThreads, 0, stream>>>(in, out, n, totals, op);
    return hipGetLastError();
}

template <class K>
hipError_t sort(void* temp, size_t& temp_bytes, const K* keys_in, K* keys_out, const uint32_t* values_in,
                uint32_t* values_out, uint32_t n, int begin_bit, int end_bit, hipStream_t stream) {
    const uint32_t tiles = std::max(1u, (n + kTile - 1) / kTile);
    const uint32_t entries = kDigits * tiles;
    size_t scan_bytes = 0;
    (void)scan<uint32_t, SumOp<uint32_t>, false>(nullptr, scan_bytes, nullptr, nullptr, entries, stream, {});
    const size_t key_bytes = align_up(size_t(n) * sizeof(K)), value_bytes = align_up(size_t(n) * 4);
    const size_t count_bytes = align_up(size_t(entries) * 4);
    const size_t need = key_bytes + value_bytes + count_bytes + scan_bytes;
    if (!temp) {
        temp_bytes = need;
        return hipSuccess;
    }
    if (temp_bytes < need) return hipErrorInvalidValue;
    if (!n) return hipSuccess;
    char* p = static_cast<char*>(temp);
    K* keys_tmp = reinterpret_cast<K*>(p);
    uint32_t* values_tmp = reinterpret_cast<uint32_t*>(p + key_bytes);
    uint32_t* counts = reinterpret_cast<uint32_t*>(p + key_bytes + value_bytes);
    void* scan_temp = p + key_bytes + value_bytes + count_bytes;
    const bool with_values = values_out != nullptr;
    const int passes = std::max(0, (end_bit - begin_bit + 7) / 8);
    if (!passes) {
        hipError_t e = hipMemcpyAsync(keys_out, keys_in, size_t(n) * sizeof(K), hipMemcpyDeviceToDevice, stream);
        if (e == hipSuccess && with_values && values_in)
            e = hipMemcpyAsync(values_out, values_in, size_t(n) * 4, hipMemcpyDeviceToDevice, stream);
        return e;
    }
    const K* src_keys = keys_in;
    const uint32_t* src_values = values_in;
    for (int pass = 0; pass < passes; ++pass) {
        // The last pass writes the output buffers.
        const bool to_out = (passes - 1 - pass) % 2 == 0;
        K* dst_keys = to_out ? keys_out : keys_tmp;
        uint32_t* dst_values = with_values ? (to_out ? values_out : values_tmp) : nullptr;
        const int shift = begin_bit + 8 * pass;
        digit_count_kernel<K><<<tiles, kThreads, 0, stream>>>(src_keys, n, shift, tiles, counts);
        size_t sb = scan_bytes;
        const hipError_t e = scan<uint32_t, SumOp<uint32_t>, false>(scan_temp, sb, counts, counts, entries, stream, {});
        if (e != hipSuccess) return e;
        // Without input values the first pass writes the input positions.
        digit_scatter_kernel<K><<<tiles, kThreads, 0, stream>>>(src_keys, dst_keys, pass == 0 ? values_in : src_values,
                                                                dst_values, n, shift, tiles, counts);
        src_keys = dst_keys;
        src_values = dst_values;
    }
    return hipGetLastError();
}

// Order-preserving encoding of a float (-0 read as +0), inverted so that an
// ascending sort is descending in the float.
__global__ void float_key_kernel(const float* in, uint32_t* out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t u = __float_as_uint(in[i]);
    if (u == 0x80000000u) u = 0;
    const uint32_t ordered = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    out[i] = ~ordered;
}

__global__ void float_unkey_kernel(const uint32_t* in, float* out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t ordered = ~in[i];
    out[i] = __uint_as_float((ordered & 0x80000000u) ? (ordered & 0x7FFFFFFFu) : ~ordered);
}

__global__ void run_heads_kernel(const uint64_t* keys, uint32_t n, uint32_t* heads) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) heads[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// slot[i]: exclusive count of heads before i; heads are read again from keys.
__global__ void run_emit_kernel(const uint64_t* keys, uint32_t n, const uint32_t* slot, uint64_t* unique,
                                uint32_t* starts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !(i == 0 || keys[i] != keys[i - 1])) return;
    unique[slot[i]] = keys[i];
    starts[slot[i]] = i;
}

__global__ void no_runs_kernel(uint32_t* runs) { *runs = 0; }

__global__ void run_count_kernel(const uint64_t* keys, uint32_t n, const uint32_t* slot, const uint32_t* starts,
                                 uint32_t* counts, uint32_t* runs) {
    const uint32_t total = slot[n - 1] + ((n == 1 || keys[n - 1] != keys[n - 2]) ? 1u : 0u);
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r == 0) *runs = total;
    if (r < total) counts[r] = (r + 1 < total ? starts[r + 1] : n) - starts[r];
}

}  // namespace

hipError_t radix_sort_pairs(void* temp, size_t& temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                            const uint32_t* values_in, uint32_t* values_out, uint32_t n, int begin_bit, int end_bit,
                            hipStream_t stream) {
    return sort<uint32_t>(temp, temp_bytes, keys_in, keys_out, values_in, values_out, n, begin_bit, end_bit, stream);
}

hipError_t radix_sort_pairs(void* temp, size_t& temp_bytes, const uint64_t* keys_in, uint64_t* keys_out,
                            const uint32_t* values_in, uint32_t* values_out, uint32_t n, int begin_bit, int end_bit,
                            hipStream_t stream) {
    return sort<uint64_t>(temp, temp_bytes, keys_in, keys_out, values_in, values_out, n, begin_bit, end_bit, stream);
}

hipError_t radix_sort_pairs_descending(void* temp, size_t& temp_bytes, const float* keys_in, float* keys_out,
                                       const uint32_t* values_in, uint32_t* values_out, uint32_t n,
                                       hipStream_t stream) {
    size_t sort_bytes = 0;
    (void)sort<uint32_t>(nullptr, sort_bytes, nullptr, nullptr, nullptr, nullptr, n, 0, 32, stream);
    const size_t key_bytes = align_up(size_t(n) * 4);
    const size_t need = 2 * key_bytes + sort_bytes;
    if (!temp) {
        temp_bytes = need;
        return hipSuccess;
    }
    if (temp_bytes < need) return hipErrorInvalidValue;
    if (!n) return hipSuccess;
    char* p = static_cast<char*>(temp);
    uint32_t* coded = reinterpret_cast<uint32_t*>(p);
    uint32_t* sorted = reinterpret_cast<uint32_t*>(p + key_bytes);
    const uint32_t blocks = (n + 255) / 256;
    float_key_kernel<<<blocks, 256, 0, stream>>>(keys_in, coded, n);
    size_t sb = sort_bytes;
    const hipError_t e = sort<uint32_t>(p + 2 * key_bytes, sb, coded, sorted, values_in, values_out, n, 0, 32, stream);
    if (e != hipSuccess) return e;
    if (keys_out) float_unkey_kernel<<<blocks, 256, 0, stream>>>(sorted, keys_out, n);
    return hipGetLastError();
}

hipError_t exclusive_sum(void* temp, size_t& temp_bytes, const uint32_t* in, uint32_t* out, uint32_t n,
                         hipStream_t stream) {
    return scan<uint32_t, SumOp<uint32_t>, false>(temp, temp_bytes, in, out, n, stream, {});
}

hipError_t exclusive_sum(void* temp, size_t& temp_bytes, const uint64_t* in, uint64_t* out, uint32_t n,
                         hipStream_t stream) {
    return scan<uint64_t, SumOp<uint64_t>, false>(temp, temp_bytes, in, out, n, stream, {});
}

hipError_t inclusive_max(void* temp, size_t& temp_bytes, const uint32_t* in, uint32_t* out, uint32_t n,
                         hipStream_t stream) {
    return scan<uint32_t, MaxOp<uint32_t>, true>(temp, temp_bytes, in, out, n, stream, {});
}

hipError_t run_length_encode(void* temp, size_t& temp_bytes, const uint64_t* keys, uint64_t* unique, uint32_t* counts,
                             uint32_t* runs, uint32_t n, hipStream_t stream) {
    size_t scan_bytes = 0;
    (void)scan<uint32_t, SumOp<uint32_t>, false>(nullptr, scan_bytes, nullptr, nullptr, n, stream, {});
    const size_t slot_bytes = align_up(size_t(n) * 4);
    const size_t need = 2 * slot_bytes + scan_bytes;
    if (!temp) {
        temp_bytes = need;
        return hipSuccess;
    }
    if (temp_bytes < need) return hipErrorInvalidValue;
    if (!n) {
        no_runs_kernel<<<1, 1, 0, stream>>>(runs);
        return hipGetLastError();
    }
    char* p = static_cast<char*>(temp);
    uint32_t* slot = reinterpret_cast<uint32_t*>(p);
    uint32_t* starts = reinterpret_cast<uint32_t*>(p + slot_bytes);
    const uint32_t blocks = (n + 255) / 256;
    run_heads_kernel<<<blocks, 256, 0, stream>>>(keys, n, slot);
    size_t sb = scan_bytes;
    const hipError_t e = scan<uint32_t, SumOp<uint32_t>, false>(p + 2 * slot_bytes, sb, slot, slot, n, stream, {});
    if (e != hipSuccess) return e;
    run_emit_kernel<<<blocks, 256, 0, stream>>>(keys, n, slot, unique, starts);
    run_count_kernel<<<blocks, 256, 0, stream>>>(keys, n, slot, starts, counts, runs);
    return hipGetLastError();
}

}  // namespace blitz_hip
