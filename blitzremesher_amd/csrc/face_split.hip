// Face split of a reference and a candidate drawn over one vertex set.
//
// A coverage mask is the OR of its mesh's faces, and a face's samples depend
// only on its three projected corners in order and its sidedness
// (src/evaluate.cpp:144-170). A candidate built from the reference by local
// edits (a collapsed copy with reused vertices, a patched LOD) keeps most of
// the reference's faces with the same vertex triple, so over a shared vertex
// stream those faces draw the same samples into both masks. The split lists
// three face sets: candidate faces with a reference face of the same ordered
// triple and sidedness (shared: drawn once, flushed into both masks),
// reference faces no candidate face matched, and unmatched candidate faces.
// OR-ing the shared and reference-only lists gives every reference face's
// samples (a matched reference face is drawn through its shared partner), and
// likewise for the candidate, so the masks are unchanged; a reference face
// matched by two equal candidate faces is simply drawn twice. On a heightfield
// and its 25 %-collapsed copy two thirds of the candidate's faces are shared,
// and the raster sets up 28 % fewer faces.
//
// Pairing: every uploaded mesh that may be a reference holds an
// open-addressing table of its faces keyed by the ordered triple (built once,
// at upload; linear probing, at most half full, so every probe sequence ends
// at an empty slot). A slot holds the triple and the face id with its
// sidedness bit, so a probe step is one 16-byte load. Per raster, each
// candidate face probes the
// reference's table; list positions come from per-block counts and one scan
// of them, so each list keeps its mesh's face order. List entries carry the
// face's id in its own mesh (bit 31: drawn two-sided), so out-of-frame faces
// and sidedness are still reported in each mesh's own face order.
#include <hip/hip_runtime.h>

#include "audit_kernels.hpp"

namespace blitz_hip {
namespace {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kTwoSided = 0x80000000u;
// Faces per block of the count and scatter kernels: 256 threads x 4 rounds
// (about 900 blocks for the 914,000 faces of a 522,000-face reference and
// its 392,000-face candidate).
constexpr uint32_t kSplitThreads = 256;
constexpr uint32_t kSplitRounds = 4;
constexpr uint32_t kSplitBlock = kSplitThreads * kSplitRounds;

__device__ __forceinline__ uint32_t triple_hash(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t h = a * 0x9E3779B1u;
    h ^= b * 0x85EBCA77u + (h << 6) + (h >> 2);
    h ^= c * 0xC2B2AE3Du + (h << 6) + (h >> 2);
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    return h;
}

__device__ __forceinline__ uint3 triple(const uint32_t* idx, uint32_t f) {
    return make_uint3(idx[3 * size_t(f)], idx[3 * size_t(f) + 1], idx[3 * size_t(f) + 2]);
}

__device__ __forceinline__ uint32_t sided(const uint8_t* two, uint32_t f) {
    return two && two[f] ? kTwoSided : 0u;
}

__global__ __launch_bounds__(256) void table_clear_kernel(uint4* table, uint32_t slots) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < slots; i += gridDim.x * blockDim.x)
        table[i] = make_uint4(0, 0, 0, kEmpty);
}

// Slot: (triple, face id | sidedness bit); the id word is claimed first, the
// triple written after (probes run in a later launch).
__global__ __launch_bounds__(256) void table_insert_kernel(const uint32_t* __restrict__ idx,
                                                           const uint8_t* __restrict__ two, uint32_t faces,
                                                           uint4* table, uint32_t mask) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= faces) return;
    const uint3 t = triple(idx, f);
    const uint32_t id = f | sided(two, f);
    uint32_t h = triple_hash(t.x, t.y, t.z) & mask;
    // At most `faces` of the > 2 * faces slots are ever taken, so a free slot
    // is reached within `faces` steps.
    for (uint32_t step = 0; step <= mask; ++step) {
        if (atomicCAS(&table[h].w, kEmpty, id) == kEmpty) {
            table[h].x = t.x;
            table[h].y = t.y;
            table[h].z = t.z;
            return;
        }
        h = (h + 1) & mask;
    }
}

// match[g]: the first reference face (in probe order) with candidate face g's
// triple and sidedness, or kEmpty; matched[f] = 1 for every reference face
// some candidate face matched (plain stores of one value).
__global__ __launch_bounds__(256) void split_probe_kernel(const uint4* __restrict__ table, uint32_t mask,
                                                          const uint32_t* __restrict__ b_idx,
                                                          const uint8_t* __restrict__ b_two, uint32_t b_faces,
                                                          uint32_t* match, uint8_t* matched) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= b_faces) return;
    const uint3 t = triple(b_idx, g);
    const uint32_t side = sided(b_two, g);
    uint32_t h = triple_hash(t.x, t.y, t.z) & mask, found = kEmpty;
    for (uint32_t step = 0; step <= mask; ++step) {
        const uint4 e = table[h];
        if (e.w == kEmpty) break;
        if (e.x == t.x && e.y == t.y && e.z == t.z && (e.w & kTwoSided) == side) {
            found = e.w & ~kTwoSided;
            break;
        }
        h = (h + 1) & mask;
    }
    match[g] = found;
    if (found != kEmpty) matched[found] = 1;
}

// Blocks [0, b_blocks) cover the candidate's faces (lists 0 = shared and
// 2 = candidate only), the rest the reference's (list 1 = reference only).
struct SplitSides {
    uint32_t a_faces, b_faces, b_blocks;
    const uint32_t* match;
    const uint8_t* matched;
};

// Which list face i of a block's side goes to (0, 1, 2), or 3 for none.
__device__ __forceinline__ uint32_t list_of(const SplitSides& s, bool b_side, uint32_t i) {
    if (b_side) return i < s.b_faces ? (s.match[i] != kEmpty ? 0u : 2u) : 3u;
    return i < s.a_faces && !s.matched[i] ? 1u : 3u;
}

// Per block and list: the number of its faces in that list.
__global__ __launch_bounds__(kSplitThreads) void split_count_kernel(SplitSides s, uint32_t* block_counts) {
    const bool b_side = blockIdx.x < s.b_blocks;
    const uint32_t blk = b_side ? blockIdx.x : blockIdx.x - s.b_blocks;
    uint32_t n0 = 0, n1 = 0, n2 = 0;
    for (uint32_t r = 0; r < kSplitRounds; ++r) {
        const uint32_t l = list_of(s, b_side, blk * kSplitBlock + r * kSplitThreads + threadIdx.x);
        n0 += l == 0;
        n1 += l == 1;
        n2 += l == 2;
    }
    for (int o = 32; o > 0; o >>= 1) {
        n0 += __shfl_xor(n0, o);
        n1 += __shfl_xor(n1, o);
        n2 += __shfl_xor(n2, o);
    }
    __shared__ uint32_t part[kSplitThreads / 64][3];
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[wave][0] = n0;
        part[wave][1] = n1;
        part[wave][2] = n2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        uint32_t v = 0;
        for (uint32_t w = 0; w < kSplitThreads / 64; ++w) v += part[w][threadIdx.x];
        block_counts[3 * size_t(blockIdx.x) + threadIdx.x] = v;
    }
}

// Exclusive prefix over blocks of every list's counts (one workgroup), and
// the list lengths.
__global__ __launch_bounds__(1024) void split_scan_kernel(uint32_t* block_counts, uint32_t blocks,
                                                          uint32_t* counts) {
    __shared__ uint32_t carry[3];
    __shared__ uint32_t wave_sum[16][3];
    if (threadIdx.x < 3) carry[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t base = 0; base < blocks; base += 1024) {
        const uint32_t b = base + threadIdx.x;
        uint32_t v[3], inc[3];
        for (int k = 0; k < 3; ++k) {
            v[k] = b < blocks ? block_counts[3 * size_t(b) + k] : 0u;
            inc[k] = v[k];
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t up = __shfl_up(inc[k], o);
                if (lane >= uint32_t(o)) inc[k] += up;
            }
        }
        if (lane == 63)
            for (int k = 0; k < 3; ++k) wave_sum[wave][k] = inc[k];
        __syncthreads();
        for (int k = 0; k < 3; ++k) {
            uint32_t before = carry[k];
            for (uint32_t w = 0; w < wave; ++w) before += wave_sum[w][k];
            if (b < blocks) block_counts[3 * size_t(b) + k] = before + inc[k] - v[k];
        }
        __syncthreads();
        if (threadIdx.x < 3) {
            uint32_t total = 0;
            for (uint32_t w = 0; w < 16; ++w) total += wave_sum[w][threadIdx.x];
            carry[threadIdx.x] += total;
        }
        __syncthreads();
    }
    if (threadIdx.x < 3) counts[threadIdx.x] = carry[threadIdx.x];
}

// Writes each block's faces at its offsets, in face order within the block.
__global__ __launch_bounds__(kSplitThreads) void split_scatter_kernel(SplitSides s, const uint32_t* a_idx,
                                                                      const uint8_t* a_two, const uint32_t* b_idx,
                                                                      const uint8_t* b_two,
                                                                      const uint32_t* block_offsets, FaceSplit out) {
    const bool b_side = blockIdx.x < s.b_blocks;
    const uint32_t blk = b_side ? blockIdx.x : blockIdx.x - s.b_blocks;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __shared__ uint32_t base[3];
    __shared__ uint32_t wave_count[kSplitThreads / 64][3];
    if (threadIdx.x < 3) base[threadIdx.x] = block_offsets[3 * size_t(blockIdx.x) + threadIdx.x];
    for (uint32_t r = 0; r < kSplitRounds; ++r) {
        const uint32_t i = blk * kSplitBlock + r * kSplitThreads + threadIdx.x;
        const uint32_t l = list_of(s, b_side, i);
        uint64_t ballot[3];
        for (uint32_t k = 0; k < 3; ++k) ballot[k] = __ballot(l == k);
        if (lane == 0)
            for (uint32_t k = 0; k < 3; ++k) wave_count[wave][k] = uint32_t(__popcll(ballot[k]));
        __syncthreads();
        if (l < 3) {
            uint32_t at = base[l] + uint32_t(__popcll(ballot[l] & ((1ull << lane) - 1)));
            for (uint32_t w = 0; w < wave; ++w) at += wave_count[w][l];
            const bool from_b = l != 1;
            const uint3 t = triple(from_b ? b_idx : a_idx, i);
            const uint32_t id = i | sided(from_b ? b_two : a_two, i);
            uint32_t* tri = l == 0 ? out.shared_tri : l == 1 ? out.a_tri : out.b_tri;
            tri[3 * size_t(at)] = t.x;
            tri[3 * size_t(at) + 1] = t.y;
            tri[3 * size_t(at) + 2] = t.z;
            if (l == 0) {
                out.shared_ids_a[at] = s.match[i] | (id & kTwoSided);
                out.shared_ids_b[at] = id;
            } else {
                (l == 1 ? out.a_ids : out.b_ids)[at] = id;
            }
        }
        __syncthreads();
        if (threadIdx.x < 3) {
            uint32_t total = 0;
            for (uint32_t w = 0; w < kSplitThreads / 64; ++w) total += wave_count[w][threadIdx.x];
            base[threadIdx.x] += total;
        }
        __syncthreads();
    }
}

}  // namespace

uint32_t face_table_slots(uint32_t faces) {
    uint32_t slots = 1024;
    while (slots < 2u * faces + 1u) slots <<= 1;
    return slots;
}

uint32_t face_split_blocks(uint32_t a_faces, uint32_t b_faces) {
    return (a_faces + kSplitBlock - 1) / kSplitBlock + (b_faces + kSplitBlock - 1) / kSplitBlock;
}

void launch_face_table(const uint32_t* idx, const uint8_t* two, uint32_t faces, uint4* table, hipStream_t stream) {
    const uint32_t slots = face_table_slots(faces);
    table_clear_kernel<<<std::min<uint32_t>(2048, (slots + 255) / 256), 256, 0, stream>>>(table, slots);
    if (faces) table_insert_kernel<<<(faces + 255) / 256, 256, 0, stream>>>(idx, two, faces, table, slots - 1);
}

void launch_face_split(const uint32_t* a_idx, const uint8_t* a_two, uint32_t a_faces, const uint4* a_table,
                       const uint32_t* b_idx, const uint8_t* b_two, uint32_t b_faces, uint32_t* match,
                       uint8_t* matched, uint32_t* block_counts, const FaceSplit& out, hipStream_t stream) {
    const uint32_t mask = face_table_slots(a_faces) - 1;
    if (a_faces) launch_fill(reinterpret_cast<uint64_t*>(matched), (a_faces + 7) / 8, 0, stream);
    if (b_faces)
        split_probe_kernel<<<(b_faces + 255) / 256, 256, 0, stream>>>(a_table, mask, b_idx, b_two, b_faces, match,
                                                                       matched);
    SplitSides s{a_faces, b_faces, (b_faces + kSplitBlock - 1) / kSplitBlock, match, matched};
    const uint32_t blocks = face_split_blocks(a_faces, b_faces);
    if (blocks) split_count_kernel<<<blocks, kSplitThreads, 0, stream>>>(s, block_counts);
    split_scan_kernel<<<1, 1024, 0, stream>>>(block_counts, blocks, out.counts);
    if (blocks)
        split_scatter_kernel<<<blocks, kSplitThreads, 0, stream>>>(s, a_idx, a_two, b_idx, b_two, block_counts, out);
}

}  // namespace blitz_hip
