// Coverage-audit kernels for gfx950 (wave64).
//
// Bit-exactness contract: every FP64 expression below is evaluated in the same
// order as the reference CPU audit (src/evaluate.cpp:136-170) and the library
// is compiled with -ffp-contract=off, so no multiply-add is fused. FP64
// division on gfx950 (div_scale/div_fmas/div_fixup) is correctly rounded.
//
// Raster pipeline: project_kernel projects every vertex of a vertex set once
// per view (a candidate that shares its vertex stream with the reference
// shares the projection); raster_band_kernel then lets one 1024-thread
// workgroup own a band of tile rows of one view's mask in LDS, walk a chunk
// of the mesh's faces, OR coverage into LDS and flush the band once.
#include <hip/hip_runtime.h>
#include <algorithm>

#include "audit_kernels.hpp"
#include "raster_device.hpp"

namespace blitz_hip {
namespace {

using namespace device;

// src/evaluate.cpp:138-141: x = dot(p - c, right) * scale * ss + size/2;
// z = distance - dot(p - c, forward); scale = focal / z in perspective.
__global__ __launch_bounds__(256) void project_kernel(const ProjectTask* __restrict__ tasks) {
    const ProjectTask& t = tasks[blockIdx.y];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= t.vertices) return;
    const float4 p = t.positions[i];
    const double x = double(p.x) - double(t.center.x);
    const double y = double(p.y) - double(t.center.y);
    const double z0 = double(p.z) - double(t.center.z);
    const Camera& c = t.camera;
    const double along_f = x * double(c.forward.x) + y * double(c.forward.y) + z0 * double(c.forward.z);
    const double z = c.distance - along_f;
    const double scale = c.perspective ? c.focal / z : c.scale;
    const double half = double(t.size) * .5;
    const double ssd = double(t.ss);
    const double along_r = x * double(c.right.x) + y * double(c.right.y) + z0 * double(c.right.z);
    const double along_u = x * double(c.up.x) + y * double(c.up.y) + z0 * double(c.up.z);
    t.xy[i] = make_double2(along_r * scale * ssd + half, along_u * scale * ssd + half);
    if (t.z) t.z[i] = z;
}

// Faces whose bounding box is at most 4 samples wide and 8 high are
// rasterized by their own lane, a row per step (lane_bitmap_wave); other faces
// within 8x8 samples take one whole-wave pass (lane = sample, ballot =
// bitmap); for larger ones the wavefront first sorts the face's tiles into
// empty, fully covered and undecided (classify_rect), then produces each
// undecided tile's word with a ballot.

// Faces at most this many samples wide are cut into 4-column quads, wider
// ones (queued, up to kWideQueue per workgroup) into 64-sample spans (see
// raster_band_kernel).
constexpr int kQuadFaceWidth = 12;
constexpr uint32_t kWideQueue = 4096;

// Workgroups a raster launch should reach (four per CU of the 256-CU part).
constexpr uint32_t kFillBlocks = 1024;

// One block per (task, band, face chunk); the hardware dispatches block w to
// XCD w % 8. With a multiple of 8 chunks, chunk c of every task runs on XCD
// c % 8, tasks in order: each XCD's L2 keeps the index ranges of its chunks
// across all views, so the index buffers come from HBM about once per launch
// rather than once per view. Otherwise the blocks of one task land on one
// XCD, whose L2 then holds the task's projected vertices.
__global__ __launch_bounds__(kRasterThreads) void raster_band_kernel(const RasterTask* __restrict__ tasks,
                                                                     RasterGrid grid) {
    extern __shared__ unsigned long long band[];
    __shared__ uint32_t wide_queue[kWideQueue];
    __shared__ uint32_t wide_count;
    uint32_t task_id, band_id, chunk;
    if (grid.chunks % 8 == 0) {
        const uint32_t xcd = blockIdx.x % 8, slot = blockIdx.x / 8, per_band = grid.chunks / 8;
        task_id = slot / (grid.bands * per_band);
        if (task_id >= grid.tasks) return;
        const uint32_t r = slot % (grid.bands * per_band);
        band_id = r / per_band;
        chunk = (r % per_band) * 8 + xcd;
    } else {
        const uint32_t total = grid.tasks * grid.bands * grid.chunks;
        const uint32_t per_xcd = (total + 7) / 8;
        const uint32_t work = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
        if (work >= total) return;
        chunk = work % grid.chunks;
        band_id = (work / grid.chunks) % grid.bands;
        task_id = work / (grid.chunks * grid.bands);
    }
    const RasterTask& t = tasks[task_id];
    const uint32_t ty0 = band_id * grid.band_rows;
    if (ty0 >= t.tiles) return;
    const uint32_t ty1 = min(t.tiles, ty0 + grid.band_rows);
    const uint32_t words = (ty1 - ty0) * t.tiles;
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) band[i] = 0;
    if (threadIdx.x == 0) wide_count = 0;
    __syncthreads();
    const int ymin = int(ty0 * kTile), ymax = min(int(ty1 * kTile), int(t.size)) - 1;
    const uint32_t faces = task_faces(t);
    const uint32_t per = (faces + grid.chunks - 1) / grid.chunks;
    const uint32_t f0 = chunk * per, f1 = min(faces, f0 + per);
    const int lane = threadIdx.x & 63;
    const uint32_t step = blockDim.x;
    // Two-deep pipeline: corners of batch i+1 and indices of batch i+2 load
    // while batch i is rasterized. Loads past the chunk re-read its last face
    // (unconditional loads keep the pipeline free of divergent branches).
    const uint32_t last = f1 > f0 ? f1 - 1 : f0;
    const auto clamp_face = [&](uint32_t f) { return min(f, last); };
    Corners next{};
    uint3 tri_next = make_uint3(0, 0, 0);
    if (f0 < f1) {
        next = load_corners(t, load_face(t, clamp_face(f0 + threadIdx.x)), clamp_face(f0 + threadIdx.x));
        tri_next = load_face(t, clamp_face(f0 + step + threadIdx.x));
    }
    for (uint32_t base = f0; base < f1; base += step) {
        const uint32_t f = base + threadIdx.x;
        const Corners corners = next;
        next = load_corners(t, tri_next, clamp_face(f + step));
        tri_next = load_face(t, clamp_face(f + 2 * step));
        Setup s;
        int kind = 0;
        if (f < f1) {
            // The frame side is converted in the loop: hoisted out of it,
            // the converted pair stays live across the whole loop and is
            // spilled to scratch. The empty asm (t is the block's task, so
            // its size is uniform) keeps the conversion in place.
            uint32_t side = t.size;
            asm volatile("" : "+s"(side));
            kind = setup_triangle(t, corners, s, double(side));
            if (corners.behind) report_behind(t);
            if (kind == 2) report_out_of_frame(t, f);
        }
        const bool drawn = kind == 1 && s.y1 >= ymin && s.y0 <= ymax;
        const int w = drawn ? s.x1 - s.x0 + 1 : 0, h = drawn ? s.y1 - s.y0 + 1 : 0;
        // Boxes at most 4 samples wide and 8 high: the lane path, an 8x8
        // local bitmap over at most 2x2 tiles, then four words.
        const bool lane_face = drawn && w <= 4 && h <= 8;
        // Faces wider than kQuadFaceWidth samples wait in the block's queue
        // and are drawn after the face loop; with the queue full they take
        // the quad items below.
        const bool big = drawn && w > kQuadFaceWidth;
        const uint64_t bigs = __ballot(big);
        bool queued = false;
        if (bigs) {
            uint32_t first = 0;
            if (lane == 0) first = atomicAdd(&wide_count, uint32_t(__popcll(bigs)));
            first = __shfl(first, 0);
            const uint32_t slot = first + uint32_t(__popcll(bigs & ((1ull << lane) - 1)));
            queued = big && slot < kWideQueue;
            if (queued) wide_queue[slot] = f;
        }
        const bool wide = drawn && !lane_face && !queued, small_box = wide && w <= 8 && h <= 8;
        if (__ballot(lane_face)) {
            const int ya = max(s.y0, ymin), yb = min(s.y1, ymax);
            const int rows = lane_face ? yb - ya + 1 : 0;
            const int max_rows = wave_max_small(rows);
            const uint64_t local = __ballot(lane_face && w == 4) ? lane_bitmap_wave<4>(s, ya, rows, max_rows, w)
                                                                 : lane_bitmap_wave<3>(s, ya, rows, max_rows, w);
            if (lane_face) {
                uint64_t word[4];
                scatter_bitmap(local, s.x0, s.y0, word);
                // Word k goes to tile row (y0 >> 3) + (k >> 1), column
                // (x0 >> 3) + (k & 1); a nonzero word holds samples of rows
                // [ya, yb], which lie in the band.
                const uint32_t at = (uint32_t(s.y0 >> 3) - ty0) * t.tiles + uint32_t(s.x0 >> 3);
                if (word[0]) atomicOr(&band[at], static_cast<unsigned long long>(word[0]));
                if (word[1]) atomicOr(&band[at + 1], static_cast<unsigned long long>(word[1]));
                if (word[2]) atomicOr(&band[at + t.tiles], static_cast<unsigned long long>(word[2]));
                if (word[3]) atomicOr(&band[at + t.tiles + 1], static_cast<unsigned long long>(word[3]));
            }
        }
        // A few faces within 8x8 boxes: one whole-wave pass each, lane l
        // testing sample (x0 + l%8, y0 + l/8); the ballot is the face's local
        // bitmap in the lane path's layout. (Same bits as the item queue
        // below, which costs more for a handful of small faces.)
        const uint64_t wide_lanes = __ballot(wide);
        if (__popcll(wide_lanes) <= 4 && __ballot(wide && !small_box) == 0) {
            uint64_t pending = wide_lanes;
            while (pending) {
                const int src = __ffsll(static_cast<unsigned long long>(pending)) - 1;
                pending &= pending - 1;
                const Setup q = shuffle_setup(s, src);
                const int lx = lane & 7, ly = lane >> 3, x = q.x0 + lx, y = q.y0 + ly;
                const bool in = x <= q.x1 && y <= q.y1 && y >= ymin && y <= ymax && covers(q, x, y);
                const uint64_t local = __ballot(in);
                if (lane < 4 && local) {
                    uint64_t word[4];
                    scatter_bitmap(local, q.x0, q.y0, word);
                    const uint64_t mine = lane == 0 ? word[0] : lane == 1 ? word[1] : lane == 2 ? word[2] : word[3];
                    if (mine) {
                        const uint32_t row = uint32_t((q.y0 >> 3) + (lane >> 1)) - ty0;
                        atomicOr(&band[row * t.tiles + uint32_t((q.x0 >> 3) + (lane & 1))],
                                 static_cast<unsigned long long>(mine));
                    }
                }
            }
            continue;
        }
        // Wider or taller faces: their boxes are cut into items, one tile
        // row by one 4-column quad (within the band), and the wave's items
        // are dealt to its lanes 64 at a time. A lane finds the face owning
        // its item from the inclusive item counts (a fixed six-step search),
        // takes that face's setup from the owner lane, classifies the item
        // exactly (classify_rect: full, empty or undecided) and tests an
        // undecided item's samples with the face's row and column products.
        uint32_t items = 0, across = 0;
        int quad0 = 0, row0 = 0, ya_face = 0, yb_face = 0;
        if (wide) {
            ya_face = max(s.y0, ymin);
            yb_face = min(s.y1, ymax);
            quad0 = s.x0 >> 2;
            row0 = ya_face >> 3;
            across = uint32_t((s.x1 >> 2) - quad0 + 1);
            items = across * uint32_t((yb_face >> 3) - row0 + 1);
        }
        uint32_t inclusive = items;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t v = __shfl_up(inclusive, d);
            if (lane >= d) inclusive += v;
        }
        const uint32_t total = __shfl(inclusive, 63);
        for (uint32_t base = 0; base < total; base += 64) {
            const uint32_t k = base + uint32_t(lane);
            const bool live = k < total;
            const uint32_t key = live ? k : total - 1;
            int lo = 0;
            for (int half = 32; half > 0; half >>= 1)
                if (__shfl(inclusive, lo + half - 1) <= key) lo += half;
            const int owner = lo;
            const uint32_t o_end = __shfl(inclusive, owner), o_items = __shfl(items, owner);
            const uint32_t o_across = __shfl(across, owner);
            const int o_quad = __shfl(quad0, owner), o_row = __shfl(row0, owner);
            const int o_ya = __shfl(ya_face, owner), o_yb = __shfl(yb_face, owner);
            const Setup q = shuffle_setup(s, owner);
            const uint32_t j = key - (o_end - o_items);
            const int qx = o_quad + int(j % o_across), ty = o_row + int(j / o_across);
            const int xa = max(4 * qx, q.x0), xb = min(4 * qx + 3, q.x1);
            const int ya = max(8 * ty, o_ya), yb = min(8 * ty + 7, o_yb);
            if (live) {
                const int kind = classify_rect(q, xa, xb, ya, yb);
                uint64_t word = 0;
                if (kind == 1)
                    word = rect_word(xa & 7, xb & 7, ya & 7, yb & 7);
                else if (kind == 2)
                    word = quad_word(q, 4 * qx, xa, xb, ya, yb);
                if (word)
                    atomicOr(&band[uint32_t(ty - int(ty0)) * t.tiles + uint32_t(qx >> 1)],
                             static_cast<unsigned long long>(word));
            }
        }
    }
    __syncthreads();
    // Queued faces, wider than kQuadFaceWidth samples: set up again and cut
    // into items of one tile row by one 64-sample span (eight tiles), dealt
    // to the lanes of each wave as above; an undecided span's rows are
    // found as intervals (span_half_words), so the work of a row does not
    // grow with the face's width. Drawn here, after the face loop, their
    // registers do not add to the loop's. Queue entries go to the waves in
    // turn (entry i0 + lane * waves + wave), so a short queue still spreads
    // over every wave of the workgroup.
    const uint32_t queued_faces = min(wide_count, kWideQueue);
    const uint32_t waves = blockDim.x / 64, wave = threadIdx.x / 64;
    for (uint32_t i0 = 0; i0 < queued_faces; i0 += blockDim.x) {
        const uint32_t i = i0 + uint32_t(lane) * waves + wave;
        Setup s;
        bool mine = false;
        if (i < queued_faces) {
            const uint32_t f = wide_queue[i];
            mine = setup_triangle(t, load_corners(t, load_face(t, f), f), s, double(t.size)) == 1;
        }
        uint32_t items = 0, across = 0;
        int span0 = 0, row0 = 0, ya_face = 0, yb_face = 0;
        if (mine) {
            ya_face = max(s.y0, ymin);
            yb_face = min(s.y1, ymax);
            span0 = s.x0 >> 6;
            row0 = ya_face >> 3;
            across = uint32_t((s.x1 >> 6) - span0 + 1);
            items = across * uint32_t((yb_face >> 3) - row0 + 1);
        }
        uint32_t inclusive = items;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t v = __shfl_up(inclusive, d);
            if (lane >= d) inclusive += v;
        }
        const uint32_t total = __shfl(inclusive, 63);
        for (uint32_t base = 0; base < total; base += 64) {
            const uint32_t k = base + uint32_t(lane);
            const bool live = k < total;
            const uint32_t key = live ? k : total - 1;
            int lo = 0;
            for (int half = 32; half > 0; half >>= 1)
                if (__shfl(inclusive, lo + half - 1) <= key) lo += half;
            const int owner = lo;
            const uint32_t o_end = __shfl(inclusive, owner), o_items = __shfl(items, owner);
            const uint32_t o_across = __shfl(across, owner);
            const int o_span = __shfl(span0, owner), o_row = __shfl(row0, owner);
            const int o_ya = __shfl(ya_face, owner), o_yb = __shfl(yb_face, owner);
            const Setup q = shuffle_setup(s, owner);
            const uint32_t j = key - (o_end - o_items);
            const int sx = o_span + int(j % o_across), ty = o_row + int(j / o_across);
            const int xa = max(64 * sx, q.x0), xb = min(64 * sx + 63, q.x1);
            const int ya = max(8 * ty, o_ya), yb = min(8 * ty + 7, o_yb);
            // Drawn faces have finite corners inside the frame
            // (box_in_frame), so classify_rect decides or leaves undecided
            // on the samples alone.
            const int kind = live ? classify_rect(q, xa, xb, ya, yb) : 0;
            if (kind == 0) continue;
            unsigned long long* row = &band[uint32_t(ty - int(ty0)) * t.tiles + uint32_t(8 * sx)];
            // Rows 0-3 and 4-7 of the tile row in turn; tiles past the
            // span's last sample get zero words.
            for (int half = 0; half < 2; ++half) {
                const int ra = max(ya, 8 * ty + 4 * half), rb = min(yb, 8 * ty + 4 * half + 3);
                if (ra > rb) continue;
                uint32_t word[8];
                span_half_words(q, 64 * sx, xa, xb, ra, rb, kind == 1, word);
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (word[c]) atomicOr(&row[c], static_cast<unsigned long long>(word[c]) << (32 * half));
            }
        }
    }
    __syncthreads();
    uint64_t* out = t.mask + size_t(ty0) * t.tiles;
    if (grid.chunks == 1 && !t.or_flush) {
        for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) out[i] = band[i];
    } else {
        // A shared face list's band goes into both masks.
        uint64_t* out2 = t.mask2 ? t.mask2 + size_t(ty0) * t.tiles : nullptr;
        for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) {
            const unsigned long long v = band[i];
            if (v) {
                atomicOr(reinterpret_cast<unsigned long long*>(out + i), v);
                if (out2) atomicOr(reinterpret_cast<unsigned long long*>(out2 + i), v);
            }
        }
    }
}

// Sets *differs when any byte of a[0..n) and b[0..n) differs (every writer
// stores the same value).
__global__ __launch_bounds__(256) void differ_kernel(const uint8_t* __restrict__ a, const uint8_t* __restrict__ b,
                                                     size_t n, uint64_t* differs) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    bool d = false;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) d |= a[i] != b[i];
    if (d) *differs = 1;
}

__global__ void fill_kernel(uint64_t* p, size_t n, uint64_t value) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        p[i] = value;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Mask summaries of src/evaluate.cpp:198-202,276-279: popcounts of A, C,
// A^C and A|C. Padding bits are never set, so tile-major counts equal the
// reference's row-major counts.
__global__ __launch_bounds__(256) void compare_kernel(const ViewPair* __restrict__ pairs, ViewSummary* summaries,
                                                      uint32_t words_per_block) {
    const ViewPair& p = pairs[blockIdx.y];
    const size_t n = size_t(p.tiles) * p.tiles;
    const size_t begin = size_t(blockIdx.x) * words_per_block;
    if (begin >= n) return;
    const size_t end = min(n, begin + words_per_block);
    unsigned long long ca = 0, cc = 0, changed = 0, total = 0;
    for (size_t i = begin + threadIdx.x; i < end; i += blockDim.x) {
        const uint64_t a = p.a[i], c = p.c[i];
        ca += __popcll(a);
        cc += __popcll(c);
        changed += __popcll(a ^ c);
        total += __popcll(a | c);
    }
    ca = wave_sum(ca);
    cc = wave_sum(cc);
    changed = wave_sum(changed);
    total = wave_sum(total);
    // One set of atomics per block: the four waves' sums meet in LDS.
    __shared__ unsigned long long part[4][4];
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[wave][0] = ca;
        part[wave][1] = cc;
        part[wave][2] = changed;
        part[wave][3] = total;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const unsigned long long v = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] +
                                     part[3][threadIdx.x];
        ViewSummary& s = summaries[blockIdx.y];
        unsigned long long* field = threadIdx.x == 0   ? &s.count_a
                                    : threadIdx.x == 1 ? &s.count_c
                                    : threadIdx.x == 2 ? &s.changed
                                                       : &s.total;
        if (v) atomicAdd(field, v);
    }
}

// A view needs the distance transform only when the masks differ and both are
// nonempty (src/evaluate.cpp:202).
__device__ __forceinline__ bool needs_distance(const ViewSummary& s) {
    return s.changed && s.count_a && s.count_c;
}

// Views where more than 1/32 of the samples changed get a dense row pass;
// sparser views search each needed row directly in the bit mask.
__device__ __forceinline__ bool dense_rows(const ViewSummary& s, uint32_t size) {
    return needs_distance(s) && s.changed > (unsigned long long)size * size / 32;
}

// Row pass of the separable EDT: for every sample, the horizontal distance to
// the nearest site in its row (kNoSite if the row has none). One block per
// row; each thread owns one byte (8 samples) of the row and block-wide
// max/min scans give the nearest sites left and right of its byte.
__global__ __launch_bounds__(1024) void row_distance_kernel(const ViewPair* __restrict__ pairs,
                                                            const ViewSummary* __restrict__ summaries) {
    const uint32_t job = blockIdx.y;  // view*2 + direction
    const ViewPair& p = pairs[job >> 1];
    if (!p.dist_a || !dense_rows(summaries[job >> 1], p.size)) return;
    const uint64_t* mask = (job & 1) ? p.a : p.c;
    uint16_t* out = (job & 1) ? p.dist_a : p.dist_c;
    const uint32_t y = blockIdx.x;
    if (y >= p.size) return;
    const uint32_t tx = threadIdx.x;
    const bool live = tx < p.tiles;
    const uint32_t byte = live ? uint32_t((mask[size_t(y >> 3) * p.tiles + tx] >> ((y & 7) << 3)) & 0xFF) : 0u;
    __shared__ int left[1024], right[1024];
    const int base = int(tx << 3);
    left[tx] = byte ? base + 31 - __clz(byte) : -1;            // last site in this byte
    right[tx] = byte ? base + __ffs(byte) - 1 : 0x7FFFFFFF;   // first site in this byte
    __syncthreads();
    for (uint32_t o = 1; o < blockDim.x; o <<= 1) {
        const int l = tx >= o ? left[tx - o] : -1;
        const int r = tx + o < blockDim.x ? right[tx + o] : 0x7FFFFFFF;
        __syncthreads();
        left[tx] = max(left[tx], l);
        right[tx] = min(right[tx], r);
        __syncthreads();
    }
    if (!live) return;
    const int before = tx ? left[tx - 1] : -1;
    const int after = tx + 1 < blockDim.x ? right[tx + 1] : 0x7FFFFFFF;
    uint16_t d[8];
    for (int k = 0; k < 8; ++k) {
        const int x = base + k;
        // Nearest site at or left of x, at or right of x.
        const uint32_t lower = byte & ((2u << k) - 1);
        const uint32_t upper = byte & ~((1u << k) - 1);
        const int l = lower ? base + 31 - __clz(lower) : before;
        const int r = upper ? base + __ffs(upper) - 1 : after;
        int best = 0x7FFFFFFF;
        if (l >= 0) best = x - l;
        if (r != 0x7FFFFFFF) best = min(best, r - x);
        d[k] = best == 0x7FFFFFFF ? kNoSite : uint16_t(best);
    }
    uint16_t* row = out + size_t(y) * p.size + base;
    const uint32_t valid = min(8u, p.size - uint32_t(base));
    for (uint32_t k = 0; k < valid; ++k) row[k] = d[k];
}

// Column pass restricted to the samples that can set the directed Hausdorff
// maximum: every covered sample of `from` that is not covered in `to`. Each
// wavefront reads hausdorff_words() tile words (one per lane), numbers the
// wave's query samples with a prefix sum of the words' popcounts and hands
// them out 64 at a time, one per lane, so a wave with many sparse words (a
// silhouette running along a tile row) does not walk its words one after
// another. Each query's
// exact squared distance is min over rows y' of (y-y')^2 + g(y'), g = rounded
// row distance (the reference's lower envelope yields the same minimum; its
// float output is a monotone rounding of it, applied once on the host). The
// scan stops once (y-y')^2 reaches the best value or the best value cannot
// exceed the maximum already found.
// Tile words per wavefront. A silhouette band can put up to 64 queries in
// every word, and each query's scan is a chain of dependent loads, so on
// large rasters (long, wide bands) a wave takes 8 words (at most 8 rounds of
// 64 queries) and the heavy part of a view spreads over many waves; small
// rasters keep 64 words per wave (fewer, fuller waves).
__host__ __device__ constexpr uint32_t hausdorff_words(uint32_t tiles) { return tiles >= 256 ? 8u : 64u; }

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, uint32_t(__shfl_xor(int(v), o)));
    return v;
}

// Position of the j-th (0-based) set bit of `bits` (j < popcount(bits)).
__device__ __forceinline__ uint32_t nth_bit(uint64_t bits, uint32_t j) {
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t width = 32; width > 0; width >>= 1) {
        const uint64_t low = bits & ((1ull << width) - 1);
        const uint32_t c = uint32_t(__popcll(low));
        if (j >= c) {
            j -= c;
            bits >>= width;
            pos += width;
        } else {
            bits = low;
        }
    }
    return pos;
}

__global__ __launch_bounds__(256) void hausdorff_kernel(const ViewPair* __restrict__ pairs,
                                                        ViewSummary* summaries) {
    const uint32_t job = blockIdx.y;  // view*2 + direction; direction 0: A -> C
    ViewSummary& s = summaries[job >> 1];
    if (!needs_distance(s)) return;
    const ViewPair& p = pairs[job >> 1];
    const uint64_t* from = (job & 1) ? p.c : p.a;
    const uint64_t* to = (job & 1) ? p.a : p.c;
    const uint16_t* dist = (job & 1) ? p.dist_a : p.dist_c;
    uint32_t* target = (job & 1) ? &s.max_ca : &s.max_ac;
    const size_t words = size_t(p.tiles) * p.tiles;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t per_wave = hausdorff_words(p.tiles);
    const size_t first = (size_t(blockIdx.x) * 4 + (threadIdx.x >> 6)) * per_wave;
    if (first >= words) return;  // whole waves
    const size_t mine = first + lane;
    const uint64_t my_bits = lane < per_wave && mine < words ? from[mine] & ~to[mine] : 0;
    if (!__ballot(my_bits != 0)) return;  // whole waves
    // Inclusive prefix sum of the words' query counts over the wave.
    const uint32_t count = uint32_t(__popcll(my_bits));
    uint32_t inclusive = count;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = uint32_t(__shfl_up(int(inclusive), o));
        if (lane >= uint32_t(o)) inclusive += v;
    }
    const uint32_t exclusive = inclusive - count;
    const uint32_t total = uint32_t(__shfl(int(inclusive), 63));
    const uint32_t size = p.size;
    // Without row-distance planes every view searches the mask directly.
    const bool dense = p.dist_a != nullptr && dense_rows(s, size);
    // Wave-uniform copy of the maximum found so far (refreshed every round).
    uint32_t known = uint32_t(__shfl(int(__hip_atomic_load(target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)), 0));
    // Verdict-only audits: once the maximum reaches fail_sq the view fails
    // whatever the exact value, and queries stop at that lower bound.
    const uint32_t fail = p.fail_sq;
    if (known >= fail) return;  // whole waves
    uint32_t local = 0;
    for (uint32_t base = 0; base < total; base += 64) {
        const uint32_t g = base + lane;
        // Owner word: the first lane whose inclusive sum exceeds g.
        uint32_t w = 0;
#pragma unroll
        for (uint32_t step = 32; step > 0; step >>= 1) {
            const uint32_t probe = uint32_t(__shfl(int(inclusive), int(w + step - 1)));
            if (probe <= g) w += step;
        }
        w = min(w, 63u);
        const uint64_t bits = __shfl(my_bits, int(w));
        const uint32_t rank = g - uint32_t(__shfl(int(exclusive), int(w)));
        const uint32_t floor_value = max(known, local);
        if (g < total && floor_value < fail) {
            const uint32_t bit = nth_bit(bits, rank);
            const size_t i = first + w;
            const uint32_t tx = uint32_t(i % p.tiles), ty = uint32_t(i / p.tiles);
            const uint32_t x = (tx << 3) | (bit & 7), y = (ty << 3) | (bit >> 3);
            // Rounded squared row distance of row r, or kFar. Sparse views
            // search the mask only as far as a site could still improve
            // `best`; the rounding of d^2 above 2^24 is covered by the +2
            // column margin.
            auto row_term = [&](uint32_t r, uint32_t best, uint32_t kk) -> uint32_t {
                uint32_t d;
                if (dense) {
                    const uint16_t v = dist[size_t(r) * size + x];
                    d = v == kNoSite ? kFar : v;
                } else {
                    const uint32_t room = best == kFar ? kFar : best - kk;
                    const uint32_t reach = room == kFar ? size : uint32_t(sqrtf(float(room))) + 2;
                    d = nearest_in_row(to, p.tiles, x, r, reach);
                }
                return d == kFar ? kFar : kk + rounded_square(d);
            };
            uint32_t best = row_term(y, kFar, 0);
            for (uint32_t k = 1; best > floor_value && uint64_t(k) * k < best; ++k) {
                // Rows within k-1 are done: the minimum is >= min(best, k^2).
                if (uint64_t(k) * k >= fail && best >= fail) {
                    best = fail;
                    break;
                }
                const uint32_t kk = k * k;
                const bool up = y >= k, down = y + k < size;
                if (!up && !down) break;
                if (up) best = min(best, row_term(y - k, best, kk));
                if (down) best = min(best, row_term(y + k, best, kk));
            }
            if (best > floor_value && best != kFar) local = best;
        }
        // Share the round's maximum with the wave and the other waves, and
        // take theirs, so later queries stop earlier.
        local = wave_max(local);
        if (base + 64 < total) {
            if (lane == 0 && local > known) atomicMax(target, local);
            known = max(known, uint32_t(__shfl(int(__hip_atomic_load(target, __ATOMIC_RELAXED,
                                                                       __HIP_MEMORY_SCOPE_AGENT)), 0)));
            if (known >= fail) return;  // whole waves
        }
    }
    if (lane == 0 && local > known) atomicMax(target, local);
}

}  // namespace

void launch_differ(const void* a, const void* b, size_t bytes, uint64_t* differs, hipStream_t stream) {
    if (!bytes) return;
    const unsigned blocks = unsigned(std::min<size_t>(1024, (bytes + 255) / 256));
    differ_kernel<<<blocks, 256, 0, stream>>>(static_cast<const uint8_t*>(a), static_cast<const uint8_t*>(b), bytes,
                                               differs);
}

void launch_fill(uint64_t* p, size_t n, uint64_t value, hipStream_t stream) {
    if (!n) return;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 65535);
    fill_kernel<<<dim3(uint32_t(blocks)), dim3(256), 0, stream>>>(p, n, value);
}

void launch_project(const ProjectTask* tasks, uint32_t count, uint32_t max_vertices, hipStream_t stream) {
    if (!count || !max_vertices) return;
    project_kernel<<<dim3((max_vertices + 255) / 256, count), dim3(256), 0, stream>>>(tasks);
}

RasterGrid make_raster_grid(uint32_t tasks, uint32_t tiles, uint32_t max_faces) {
    RasterGrid g{};
    g.tasks = tasks;
    g.band_rows = std::max<uint32_t>(1, std::min<uint32_t>(tiles, kBandWords / std::max(1u, tiles)));
    g.bands = (tiles + g.band_rows - 1) / g.band_rows;
    // Up to 32 face chunks per task (one per CU of an XCD, so an XCD works on
    // one view at a time and its L2 holds that view's projected vertices),
    // each with at least 8 batches of 1024 faces.
    g.chunks = std::max<uint32_t>(1, std::min<uint32_t>(32, max_faces / (8 * kRasterThreads)));
    // Few views of a large raster (few tasks x bands): split the faces
    // further, down to one batch per chunk, so the launch still has about
    // four workgroups per CU.
    const uint32_t base = std::max(1u, tasks * g.bands);
    const uint32_t wanted = (kFillBlocks + base - 1) / base;
    if (g.chunks < wanted)
        g.chunks = std::max(g.chunks, std::min(wanted, (max_faces + kRasterThreads - 1) / kRasterThreads));
    // Beyond 8 chunks, a multiple of 8 (one chunk set per XCD).
    if (g.chunks > 8) g.chunks = (g.chunks + 7) / 8 * 8;
    return g;
}

void launch_raster(const RasterTask* tasks, const RasterGrid& grid, uint32_t tiles, hipStream_t stream) {
    const uint32_t total = grid.tasks * grid.bands * grid.chunks;
    if (!total) return;
    const size_t lds = size_t(grid.band_rows) * tiles * sizeof(uint64_t);
    static bool configured = false;
    if (!configured) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&raster_band_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(kBandWords * sizeof(uint64_t)));
        configured = true;
    }
    const uint32_t blocks = (total + 7) / 8 * 8;
    raster_band_kernel<<<dim3(blocks), dim3(kRasterThreads), lds, stream>>>(tasks, grid);
}

void launch_compare(const ViewPair* pairs, ViewSummary* summaries, uint32_t count, uint32_t tiles,
                    hipStream_t stream) {
    if (!count) return;
    const size_t words = size_t(tiles) * tiles;
    const uint32_t per_block = 256 * 4;
    const uint32_t blocks = uint32_t((words + per_block - 1) / per_block);
    compare_kernel<<<dim3(blocks, count), dim3(256), 0, stream>>>(pairs, summaries, per_block);
}

void launch_distance(const ViewPair* pairs, ViewSummary* summaries, uint32_t count, uint32_t size, uint32_t tiles,
                     hipStream_t stream) {
    if (!count) return;
    const uint32_t threads = std::max<uint32_t>(64, (tiles + 63) / 64 * 64);
    row_distance_kernel<<<dim3(size, count * 2), dim3(threads), 0, stream>>>(pairs, summaries);
    const size_t words = size_t(tiles) * tiles;
    const uint32_t per_block = 4 * hausdorff_words(tiles);
    hausdorff_kernel<<<dim3(uint32_t((words + per_block - 1) / per_block), count * 2), dim3(256), 0, stream>>>(
        pairs, summaries);
}

void launch_hausdorff(const ViewPair* pairs, ViewSummary* summaries, uint32_t count, uint32_t tiles,
                      hipStream_t stream) {
    if (!count) return;
    const size_t words = size_t(tiles) * tiles;
    const uint32_t per_block = 4 * hausdorff_words(tiles);
    hausdorff_kernel<<<dim3(uint32_t((words + per_block - 1) / per_block), count * 2), dim3(256), 0, stream>>>(
        pairs, summaries);
}

}  // namespace blitz_hip
