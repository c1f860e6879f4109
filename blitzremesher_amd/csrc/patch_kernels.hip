// Bound-reference patch audit kernels for gfx950.
//
// A bound reference keeps, per view, its projection, coverage mask, covered
// sample count and a bin of its drawable faces per 8x8 tile (the role of the
// reference's RasterMemo and tile bins, src/neural/evaluate.cu:173-197,
// 386-511). A patch replaces a set of reference faces by new faces over the
// same vertices. Coverage is an OR over faces, so the candidate's mask can
// differ from the reference's only inside the tiles touched by the bounding
// boxes of the removed and added faces (the dirty-tile argument of
// src/neural/evaluate.cu:216-273). patch_kernel rebuilds exactly those tiles
// for one (view, patch): OR of the kept reference faces binned there and the
// added faces, with the same FP64 setup and predicate as the full raster;
// then the mask summaries and the exact directed Hausdorff maxima, whose
// queries all lie in the dirty tiles and whose row searches read the
// rebuilt tiles there and the reference mask elsewhere.
#include <hip/hip_runtime.h>

#include <climits>
#include <stdexcept>

#include "patch_kernels.hpp"
#include "device_sort.hpp"
#include "raster_device.hpp"

namespace blitz_hip {
void hip_check(hipError_t e, const char* what);

namespace {

using namespace device;

__device__ __forceinline__ unsigned long long wave_total(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Tile box of face f of view t when it is drawable (setup kind 1).
__device__ __forceinline__ bool face_tiles(const RasterTask& t, uint32_t f, int& tx0, int& ty0, int& tx1, int& ty1) {
    const Corners k = load_corners(t, load_face(t, f), f);
    Setup s;
    if (setup_triangle(t, k, s, double(t.size)) != 1) return false;
    tx0 = s.x0 >> 3;
    ty0 = s.y0 >> 3;
    tx1 = s.x1 >> 3;
    ty1 = s.y1 >> 3;
    return true;
}

__global__ __launch_bounds__(256) void bin_count_kernel(const RasterTask* __restrict__ views,
                                                        unsigned long long* __restrict__ counts) {
    const RasterTask& t = views[blockIdx.y];
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= t.faces) return;
    int tx0, ty0, tx1, ty1;
    if (!face_tiles(t, f, tx0, ty0, tx1, ty1)) return;
    unsigned long long* c = counts + size_t(blockIdx.y) * t.tiles * t.tiles;
    for (int ty = ty0; ty <= ty1; ++ty)
        for (int tx = tx0; tx <= tx1; ++tx) atomicAdd(&c[size_t(ty) * t.tiles + tx], 1ull);
}

__global__ __launch_bounds__(256) void bin_fill_kernel(const RasterTask* __restrict__ views,
                                                       unsigned long long* __restrict__ cursor,
                                                       uint32_t* __restrict__ faces) {
    const RasterTask& t = views[blockIdx.y];
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= t.faces) return;
    int tx0, ty0, tx1, ty1;
    if (!face_tiles(t, f, tx0, ty0, tx1, ty1)) return;
    unsigned long long* c = cursor + size_t(blockIdx.y) * t.tiles * t.tiles;
    for (int ty = ty0; ty <= ty1; ++ty)
        for (int tx = tx0; tx <= tx1; ++tx) faces[atomicAdd(&c[size_t(ty) * t.tiles + tx], 1ull)] = f;
}

// Sorted-list membership in LDS.
__device__ __forceinline__ bool contains(const uint32_t* list, uint32_t n, uint32_t v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (list[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < n && list[lo] == v;
}

constexpr uint32_t kFallback = 1;

// One 256-thread block per (view, patch); blocks are view-major.
__global__ __launch_bounds__(256) void patch_kernel(PatchInput in) {
    __shared__ uint32_t removed[kPatchMaxFaces];
    __shared__ unsigned long long rect[kPatchRectWords];
    __shared__ uint32_t prefix[kPatchRectWords + 1];
    __shared__ uint32_t partial[256];
    __shared__ int box[4];
    __shared__ uint32_t flags;
    __shared__ unsigned long long sums[4];
    __shared__ uint32_t maxes[2];
    const uint32_t view = blockIdx.x / in.patches, p = blockIdx.x % in.patches;
    const uint32_t tid = threadIdx.x;
    const RasterTask vt = in.views[view];
    RasterTask at = vt;  // added faces: one sidedness for all
    at.two_sided = nullptr;
    at.force_two = vt.force_two || in.added_two_sided;
    const uint32_t tiles = in.tiles, size = vt.size;
    const uint32_t r0 = in.removed_offsets[p], nr = in.removed_offsets[p + 1] - r0;
    const uint32_t a0 = in.added_offsets[p], na = in.added_offsets[p + 1] - a0;
    for (uint32_t i = tid; i < nr; i += blockDim.x) removed[i] = in.removed[r0 + i];
    if (tid == 0) {
        box[0] = box[1] = INT_MAX;
        box[2] = box[3] = -1;
        flags = 0;
        sums[0] = sums[1] = sums[2] = sums[3] = 0;
        maxes[0] = maxes[1] = 0;
    }
    __syncthreads();
    auto added_face = [&](uint32_t i) {
        const uint32_t* q = in.added + 3 * size_t(a0 + i);
        return make_uint3(q[0], q[1], q[2]);
    };
    // Dirty rectangle: tile box of every drawable removed or added face.
    for (uint32_t i = tid; i < nr + na; i += blockDim.x) {
        const bool add = i >= nr;
        const RasterTask& t = add ? at : vt;
        const uint32_t f = add ? 0u : removed[i];
        const Corners k = load_corners(t, add ? added_face(i - nr) : load_face(vt, f), f);
        Setup s;
        const int kind = setup_triangle(t, k, s, double(t.size));
        if (kind == 2 || k.behind) atomicOr(&flags, kFallback);
        if (kind == 1) {
            atomicMin(&box[0], s.x0 >> 3);
            atomicMin(&box[1], s.y0 >> 3);
            atomicMax(&box[2], s.x1 >> 3);
            atomicMax(&box[3], s.y1 >> 3);
        }
    }
    __syncthreads();
    PatchSummary* out = in.out + size_t(blockIdx.x);
    const unsigned long long count_a = in.count_a[view];
    const int bx0 = box[0], by0 = box[1], bx1 = box[2], by1 = box[3];
    const uint32_t width = bx1 >= bx0 ? uint32_t(bx1 - bx0 + 1) : 0u;
    const uint32_t n = bx1 >= bx0 ? width * uint32_t(by1 - by0 + 1) : 0u;
    if (flags || n > kPatchRectWords || n == 0) {
        if (tid == 0) {
            PatchSummary r{};
            r.flags = (flags || n > kPatchRectWords) ? kFallback : 0u;
            r.count_c = count_a;  // no drawable face changes: the masks are equal
            r.total = count_a;
            *out = r;
        }
        return;
    }
    const uint64_t* mask_a = in.masks + size_t(view) * tiles * tiles;
    const unsigned long long* offsets = in.bin_offsets + size_t(view) * tiles * tiles;
    auto tile_of = [&](uint32_t w) { return uint32_t(by0 + int(w / width)) * tiles + uint32_t(bx0 + int(w % width)); };
    // Bin sizes of the dirty tiles, exclusive prefix in `prefix`.
    {
        uint32_t local[4], run = 0;
        for (int j = 0; j < 4; ++j) {
            const uint32_t w = tid * 4 + j;
            uint32_t c = 0;
            if (w < n) {
                const uint32_t tile = tile_of(w);
                c = uint32_t(offsets[tile + 1] - offsets[tile]);
            }
            run += c;
            local[j] = run;
        }
        partial[tid] = run;
        __syncthreads();
        for (uint32_t o = 1; o < blockDim.x; o <<= 1) {
            const uint32_t v = tid >= o ? partial[tid - o] : 0u;
            __syncthreads();
            partial[tid] += v;
            __syncthreads();
        }
        const uint32_t before = tid ? partial[tid - 1] : 0u;
        for (int j = 0; j < 4; ++j) {
            const uint32_t w = tid * 4 + j;
            if (w < n) prefix[w + 1] = before + local[j];
        }
        if (tid == 0) prefix[0] = 0;
        for (uint32_t w = tid; w < n; w += blockDim.x) rect[w] = 0;
    }
    __syncthreads();
    const int ymin = by0 * 8, ymax = min((by1 + 1) * 8, int(size)) - 1;
    // OR one face into the dirty tiles.
    auto draw = [&](const Setup& s) {
        const int ya = max(s.y0, ymin), yb = min(s.y1, ymax);
        if (ya > yb) return;
        const int w = s.x1 - s.x0 + 1, h = s.y1 - s.y0 + 1;
        if (w <= 8 && h <= 8) {
            uint64_t word[4];
            scatter_bitmap(w <= 4 ? lane_bitmap4(s, ya, yb) : lane_bitmap(s, ya, yb), s.x0, s.y0, word);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int tx = (s.x0 >> 3) + (k & 1), ty = (s.y0 >> 3) + (k >> 1);
                if (word[k] && tx >= bx0 && tx <= bx1 && ty >= by0 && ty <= by1)
                    atomicOr(&rect[uint32_t(ty - by0) * width + uint32_t(tx - bx0)],
                             static_cast<unsigned long long>(word[k]));
            }
            return;
        }
        // Larger faces: each tile of the box within the dirty rectangle is
        // classified exactly first (classify_rect); only undecided tiles
        // are tested sample by sample.
        for (int ty = max(s.y0 >> 3, by0); ty <= min(s.y1 >> 3, by1); ++ty)
            for (int tx = max(s.x0 >> 3, bx0); tx <= min(s.x1 >> 3, bx1); ++tx) {
                const int xa = max(tx * 8, s.x0), xb = min(tx * 8 + 7, s.x1);
                const int ya = max(ty * 8, s.y0), yb = min(ty * 8 + 7, s.y1);
                const int kind = classify_rect(s, xa, xb, ya, yb);
                if (kind == 0) continue;
                uint64_t word = 0;
                if (kind == 1) {
                    word = rect_word(xa & 7, xb & 7, ya & 7, yb & 7);
                } else {
                    for (int y = ya; y <= yb; ++y)
                        for (int x = xa; x <= xb; ++x)
                            if (covers(s, x, y)) word |= 1ull << ((y & 7) * 8 + (x & 7));
                }
                if (word)
                    atomicOr(&rect[uint32_t(ty - by0) * width + uint32_t(tx - bx0)],
                             static_cast<unsigned long long>(word));
            }
    };
    const uint32_t entries = prefix[n];
    for (uint32_t e = tid; e < na + entries; e += blockDim.x) {
        if (e < na) {
            const Corners k = load_corners(at, added_face(e), 0);
            Setup s;
            if (setup_triangle(at, k, s, double(at.size)) == 1) draw(s);
            continue;
        }
        const uint32_t j = e - na;
        uint32_t lo = 0, hi = n;  // last w with prefix[w] <= j
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (prefix[mid] <= j)
                lo = mid;
            else
                hi = mid;
        }
        const uint32_t tile = tile_of(lo);
        const uint32_t f = in.bin_faces[offsets[tile] + (j - prefix[lo])];
        if (contains(removed, nr, f)) continue;
        const Corners k = load_corners(vt, load_face(vt, f), f);
        Setup s;
        if (setup_triangle(vt, k, s, double(vt.size)) != 1) continue;
        // A face binned to several dirty tiles is drawn from the first one.
        const int fx = max(s.x0 >> 3, bx0), fy = max(s.y0 >> 3, by0);
        if (uint32_t(fy) * tiles + uint32_t(fx) != tile) continue;
        draw(s);
    }
    __syncthreads();
    // Summaries over the dirty tiles (the masks agree everywhere else).
    {
        unsigned long long ca = 0, cc = 0, ch = 0, extra = 0;
        for (uint32_t w = tid; w < n; w += blockDim.x) {
            const uint64_t a = mask_a[tile_of(w)], c = rect[w];
            ca += __popcll(a);
            cc += __popcll(c);
            ch += __popcll(a ^ c);
            extra += __popcll(c & ~a);
        }
        ca = wave_total(ca);
        cc = wave_total(cc);
        ch = wave_total(ch);
        extra = wave_total(extra);
        if ((tid & 63) == 0) {
            atomicAdd(&sums[0], ca);
            atomicAdd(&sums[1], cc);
            atomicAdd(&sums[2], ch);
            atomicAdd(&sums[3], extra);
        }
    }
    __syncthreads();
    const unsigned long long count_c = count_a - sums[0] + sums[1], changed = sums[2];
    if (changed && count_a && count_c) {
        // Row byte of the candidate mask: rebuilt tiles inside the rectangle.
        auto cand_row = [&](uint32_t y) {
            const int ty = int(y >> 3);
            const uint32_t shift = (y & 7) << 3;
            return [=](uint32_t t) -> uint32_t {
                const int tx = int(t);
                const uint64_t word = (ty >= by0 && ty <= by1 && tx >= bx0 && tx <= bx1)
                                          ? rect[uint32_t(ty - by0) * width + uint32_t(tx - bx0)]
                                          : mask_a[size_t(ty) * tiles + t];
                return uint32_t(word >> shift) & 0xFFu;
            };
        };
        for (uint32_t w = tid; w < n; w += blockDim.x) {
            const uint64_t a = mask_a[tile_of(w)], c = rect[w];
            const uint32_t tx = uint32_t(bx0 + int(w % width)), ty = uint32_t(by0 + int(w / width));
            for (int dir = 0; dir < 2; ++dir) {
                // dir 0: samples of A not in C, distance to C; dir 1: C not in A, to A.
                uint64_t bits = dir == 0 ? a & ~c : c & ~a;
                uint32_t local = 0;
                while (bits) {
                    const int b = __ffsll(static_cast<unsigned long long>(bits)) - 1;
                    bits &= bits - 1;
                    const uint32_t x = (tx << 3) | uint32_t(b & 7), y = (ty << 3) | uint32_t(b >> 3);
                    auto row_term = [&](uint32_t r, uint32_t best, uint32_t kk) -> uint32_t {
                        const uint32_t room = best == kFar ? kFar : best - kk;
                        const uint32_t reach = room == kFar ? size : uint32_t(sqrtf(float(room))) + 2;
                        const uint32_t d = dir == 0 ? nearest_in_row_by(cand_row(r), tiles, x, reach)
                                                    : nearest_in_row(mask_a, tiles, x, r, reach);
                        return d == kFar ? kFar : kk + rounded_square(d);
                    };
                    const uint32_t floor_value =
                        max(local, __hip_atomic_load(&maxes[dir], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
                    uint32_t best = row_term(y, kFar, 0);
                    for (uint32_t k = 1; best > floor_value && uint64_t(k) * k < best; ++k) {
                        const uint32_t kk = k * k;
                        const bool up = y >= k, down = y + k < size;
                        if (!up && !down) break;
                        if (up) best = min(best, row_term(y - k, best, kk));
                        if (down) best = min(best, row_term(y + k, best, kk));
                    }
                    if (best > floor_value && best != kFar) local = best;
                }
                if (local) atomicMax(&maxes[dir], local);
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        PatchSummary r{};
        r.count_c = count_c;
        r.changed = changed;
        r.total = count_a + sums[3];
        r.max_ac = maxes[0];
        r.max_ca = maxes[1];
        *out = r;
    }
}

}  // namespace

void launch_bin_count(const RasterTask* views, uint32_t count, uint32_t max_faces, unsigned long long* counts,
                      hipStream_t stream) {
    if (!count || !max_faces) return;
    bin_count_kernel<<<dim3((max_faces + 255) / 256, count), dim3(256), 0, stream>>>(views, counts);
}

void launch_bin_fill(const RasterTask* views, uint32_t count, uint32_t max_faces, unsigned long long* cursor,
                     uint32_t* faces, hipStream_t stream) {
    if (!count || !max_faces) return;
    bin_fill_kernel<<<dim3((max_faces + 255) / 256, count), dim3(256), 0, stream>>>(views, cursor, faces);
}

size_t scan_workspace(size_t n) {
    if (n > UINT32_MAX) throw std::invalid_argument("scan exceeds 32-bit counts");
    size_t bytes = 0;
    hip_check(exclusive_sum(nullptr, bytes, static_cast<const uint64_t*>(nullptr), nullptr, uint32_t(n), nullptr),
              "scan workspace");
    return bytes;
}

void launch_exclusive_scan(void* temp, size_t temp_bytes, const unsigned long long* in, unsigned long long* out,
                           size_t n, hipStream_t stream) {
    static_assert(sizeof(unsigned long long) == sizeof(uint64_t));
    hip_check(exclusive_sum(temp, temp_bytes, reinterpret_cast<const uint64_t*>(in), reinterpret_cast<uint64_t*>(out),
                            uint32_t(n), stream),
              "exclusive scan");
}

void launch_patches(const PatchInput& in, hipStream_t stream) {
    const size_t blocks = size_t(in.view_count) * in.patches;
    if (!blocks) return;
    patch_kernel<<<dim3(uint32_t(blocks)), dim3(256), 0, stream>>>(in);
}

}  // namespace blitz_hip
