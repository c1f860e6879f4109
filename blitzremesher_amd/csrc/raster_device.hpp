// Device-side building blocks shared by the audit kernels (audit_kernels.hip)
// and the patch kernels (patch_kernels.hip): the reference's FP64 triangle
// setup and conservative coverage predicate, the lane bitmap path, and the
// row search of the exact distance transform.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "audit_types.hpp"

namespace blitz_hip {
namespace device {

// src/evaluate.cpp:21
__device__ __forceinline__ double edge(double ax, double ay, double bx, double by, double x, double y) {
    return (bx - ax) * (y - ay) - (by - ay) * (x - ax);
}

// std::min / std::max over an initializer list: first smallest (largest)
// element under operator<, which matters only for NaN inputs.
__device__ __forceinline__ double min3(double a, double b, double c) {
    double m = a;
    if (b < m) m = b;
    if (c < m) m = c;
    return m;
}
__device__ __forceinline__ double max3(double a, double b, double c) {
    double m = a;
    if (m < b) m = b;
    if (m < c) m = c;
    return m;
}

struct Setup {
    double ax, ay, bx, by, cx, cy, r0, r1, r2;
    int x0, x1, y0, y1;
};

// Pointers read from task descriptors are generic; loads through them would
// be flat loads, which wait on LDS and memory counters alike. Every buffer
// here is device memory, so they are read through global pointers.
template <class T> using global_ptr = const __attribute__((address_space(1))) T*;
template <class T> __device__ __forceinline__ global_ptr<T> global(const T* p) {
    return (global_ptr<T>)p;
}

// Everything one face's setup reads: its projected corners and flags. Loaded
// one batch ahead of use (the indices two batches ahead) so the gathers of the
// next batch are in flight while the current one is rasterized.
struct Corners {
    double2 a, b, c;
    bool behind, two;  // a corner at z <= 0 (perspective); face drawn two-sided
};

__device__ __forceinline__ uint3 load_face(const RasterTask& t, uint32_t f) {
    const global_ptr<uint32_t> tri = global(t.indices) + 3 * size_t(f);
    return make_uint3(tri[0], tri[1], tri[2]);
}

__device__ __forceinline__ Corners load_corners(const RasterTask& t, uint3 tri, uint32_t f) {
    Corners k;
    k.behind = false;
    if (t.z) {
        const global_ptr<double> z = global(t.z);
        k.behind = (z[tri.x] <= 0) | (z[tri.y] <= 0) | (z[tri.z] <= 0);
    }
    const global_ptr<double> xy = global(reinterpret_cast<const double*>(t.xy));
    k.a = make_double2(xy[2 * size_t(tri.x)], xy[2 * size_t(tri.x) + 1]);
    k.b = make_double2(xy[2 * size_t(tri.y)], xy[2 * size_t(tri.y) + 1]);
    k.c = make_double2(xy[2 * size_t(tri.z)], xy[2 * size_t(tri.z) + 1]);
    k.two = t.force_two || (t.ids ? (global(t.ids)[f] >> 31) != 0u : (t.two_sided && global(t.two_sided)[f]));
    return k;
}

// Face f's index in its own mesh (split face lists carry it in `ids`).
__device__ __forceinline__ uint32_t mesh_face(const uint32_t* ids, uint32_t f) {
    return ids ? global(ids)[f] & 0x7FFFFFFFu : f;
}

// Faces a task draws: its device-side list length when it has one.
__device__ __forceinline__ uint32_t task_faces(const RasterTask& t) {
    return t.count ? min(*global(t.count), t.faces) : t.faces;
}

// Status words of a face's task (and of the second mesh of a shared face
// list): the first out-of-frame face in mesh order, the behind-camera flag.
__device__ __forceinline__ void report_out_of_frame(const RasterTask& t, uint32_t f) {
    atomicMin(&t.status[0], mesh_face(t.ids, f));
    if (t.status2) atomicMin(&t.status2[0], mesh_face(t.ids2, f));
}
__device__ __forceinline__ void report_behind(const RasterTask& t) {
    t.status[1] = 1u;
    if (t.status2) t.status2[1] = 1u;
}

// Bounding box and frame test of src/evaluate.cpp:155-160 for corners a, d, e
// (d and e already swapped for a back face) whose signed area is `area`:
// false when the face is out of frame. The reference takes std::min /
// std::max over the corners in order and then tests isfinite and the frame.
// Without a NaN coordinate the order does not matter (results can differ
// only in the sign of a zero, which neither floor() nor the frame test
// sees) and the isfinite tests are implied by the frame test (an infinite
// minimum or maximum fails it), so min/max instructions do. A NaN
// coordinate makes the signed area NaN (and leaves the corners unswapped);
// such faces, and infinite ones whose area is NaN, take the reference's
// ordered form.
__device__ __forceinline__ bool box_in_frame(double2 a, double2 d, double2 e, double area, double size,
                                             double& minx, double& maxx, double& miny, double& maxy) {
    if (!isnan(area)) {
        minx = fmin(fmin(a.x, d.x), e.x);
        maxx = fmax(fmax(a.x, d.x), e.x);
        miny = fmin(fmin(a.y, d.y), e.y);
        maxy = fmax(fmax(a.y, d.y), e.y);
        return !(minx < 0 || miny < 0 || maxx > size || maxy > size);
    }
    minx = min3(a.x, d.x, e.x);
    maxx = max3(a.x, d.x, e.x);
    miny = min3(a.y, d.y, e.y);
    maxy = max3(a.y, d.y, e.y);
    return !(!isfinite(minx) || !isfinite(maxx) || !isfinite(miny) || !isfinite(maxy) || minx < 0 || miny < 0 ||
             maxx > size || maxy > size);
}

// Triangle setup of src/evaluate.cpp:144-165 on projected vertices.
// Returns 0 = nothing to draw, 1 = draw, 2 = out of frame (the reference stops
// at the first such face).
// `side` is double(t.size), the frame's side in samples.
__device__ __forceinline__ int setup_triangle(const RasterTask& t, const Corners& k, Setup& s, double side) {
    if (k.behind) return 0;
    const double2 a = k.a;
    double2 d = k.b, e = k.c;
    const double area = edge(a.x, a.y, d.x, d.y, e.x, e.y);
    if (fabs(area) < 1e-16) return 0;
    const bool back = area < 0;
    if (back && !k.two) return 0;
    if (back) {
        const double2 tmp = d;
        d = e;
        e = tmp;
    }
    double minx, maxx, miny, maxy;
    if (!box_in_frame(a, d, e, area, side, minx, maxx, miny, maxy)) return 2;
    s.x0 = max(0, int(floor(minx)));
    s.x1 = min(int(t.size) - 1, int(floor(maxx)));
    s.y0 = max(0, int(floor(miny)));
    s.y1 = min(int(t.size) - 1, int(floor(maxy)));
    s.r0 = .5 * (fabs(e.x - d.x) + fabs(e.y - d.y));
    s.r1 = .5 * (fabs(a.x - e.x) + fabs(a.y - e.y));
    s.r2 = .5 * (fabs(d.x - a.x) + fabs(d.y - a.y));
    s.ax = a.x;
    s.ay = a.y;
    s.bx = d.x;
    s.by = d.y;
    s.cx = e.x;
    s.cy = e.y;
    return 1;
}

// Conservative coverage predicate at sample (x, y) (src/evaluate.cpp:166-168).
// The reference's test fl(u + r) < 0 is evaluated as u < -r: the rounded sum
// of two doubles has the sign of the exact sum (a nonzero sum too small to
// round away is exact, by Sterbenz's lemma), overflow keeps the sign, and
// NaN fails both forms; -r is exact.
__device__ __forceinline__ bool covers(const Setup& s, int x, int y) {
    const double px = x + .5, py = y + .5;
    const double u = edge(s.bx, s.by, s.cx, s.cy, px, py);
    const double v = edge(s.cx, s.cy, s.ax, s.ay, px, py);
    const double w = edge(s.ax, s.ay, s.bx, s.by, px, py);
    return !(u < -s.r0 || v < -s.r1 || w < -s.r2);
}

// Whole-rectangle form of covers() over the samples [xa, xb] x [ya, yb]:
// 0 = no sample covered, 1 = every sample covered, 2 = undecided (test each
// sample). Exact, with no tolerance: an edge value is
// fl(fl(A*fl(py-Py)) - fl(B*fl(px-Px))) and every rounded operation is
// monotone in each operand, so over a rectangle the rounded value is
// smallest at the row end chosen by sign(A) combined with the column end
// chosen by sign(B), and largest at the opposite pair, both of which are
// samples of the rectangle. Non-finite corners are left undecided.
__device__ __forceinline__ int classify_rect(const Setup& s, int xa, int xb, int ya, int yb) {
    const double pxa = xa + .5, pxb = xb + .5, pya = ya + .5, pyb = yb + .5;
    bool all = true, none = false;
    const auto edge_range = [&](double ax, double ay, double bx, double by, double r) {
        const double a = bx - ax, b = by - ay;  // value = a*(py-ay) - b*(px-ax)
        const double t1a = a * (pya - ay), t1b = a * (pyb - ay);
        const double t2a = b * (pxa - ax), t2b = b * (pxb - ax);
        const double t1min = a >= 0 ? t1a : t1b, t1max = a >= 0 ? t1b : t1a;
        const double t2min = b >= 0 ? t2a : t2b, t2max = b >= 0 ? t2b : t2a;
        // fl(value + r) < 0 tested as value < -r (see covers()).
        const double lo = t1min - t2max, hi = t1max - t2min;
        all = all && !(lo < -r);
        none = none || hi < -r;
    };
    edge_range(s.bx, s.by, s.cx, s.cy, s.r0);
    edge_range(s.cx, s.cy, s.ax, s.ay, s.r1);
    edge_range(s.ax, s.ay, s.bx, s.by, s.r2);
    const bool finite = isfinite(s.ax) && isfinite(s.ay) && isfinite(s.bx) && isfinite(s.by) && isfinite(s.cx) &&
                        isfinite(s.cy);
    if (!finite) return 2;
    return none ? 0 : all ? 1 : 2;
}

// Bits of the samples [xa, xb] x [ya, yb] of one 8x8 tile word (coordinates
// within the tile, bit 8*y + x).
__device__ __forceinline__ uint64_t rect_word(int xa, int xb, int ya, int yb) {
    const uint64_t row = uint64_t((0xFFu >> (7 - xb)) & (0xFFu << xa));
    const uint64_t rows = (~0ull >> (8 * (7 - yb))) & (~0ull << (8 * ya));
    return (row * 0x0101010101010101ull) & rows;
}

__device__ __forceinline__ Setup shuffle_setup(const Setup& s, int lane) {
    Setup o;
    o.ax = __shfl(s.ax, lane);
    o.ay = __shfl(s.ay, lane);
    o.bx = __shfl(s.bx, lane);
    o.by = __shfl(s.by, lane);
    o.cx = __shfl(s.cx, lane);
    o.cy = __shfl(s.cy, lane);
    o.r0 = __shfl(s.r0, lane);
    o.r1 = __shfl(s.r1, lane);
    o.r2 = __shfl(s.r2, lane);
    o.x0 = __shfl(s.x0, lane);
    o.x1 = __shfl(s.x1, lane);
    o.y0 = __shfl(s.y0, lane);
    o.y1 = __shfl(s.y1, lane);
    return o;
}

// Lane path: walks the bounding box rows [ya, yb] two samples of a row per
// step, in one flattened loop (lanes of a wave finish together after
// max(rows * ceil(w/2)) steps), and collects coverage in an 8x8 bitmap
// relative to (x0, y0), bit 8*(y-y0) + (x-x0). The y-dependent products are
// recomputed every step, which costs less than a divergent row change. Every
// value is the reference expression's own rounding:
// u = (cx-bx)*(py-by) - (cy-by)*(px-bx), covered iff !(u+r0 < 0 || ...).
__device__ __forceinline__ uint64_t lane_bitmap(const Setup& s, int ya, int yb) {
    const double au = s.cx - s.bx, bu = s.cy - s.by;
    const double av = s.ax - s.cx, bv = s.ay - s.cy;
    const double aw = s.bx - s.ax, bw = s.by - s.ay;
    const int n = ((s.x1 - s.x0 + 2) >> 1) * (yb - ya + 1);
    int x = s.x0, y = ya;
    uint64_t bits = 0;
    for (int k = 0; k < n; ++k) {
        const double py = double(y) + .5;
        const double ru = au * (py - s.by), rv = av * (py - s.cy), rw = aw * (py - s.ay);
        const double px = double(x) + .5, qx = px + 1.0;
        // Bitwise | so that all three tests issue without branches.
        // fl(u + r) < 0 tested as u < -r (see covers()).
        const bool in0 = !(((ru - bu * (px - s.bx)) < -s.r0) | ((rv - bv * (px - s.cx)) < -s.r1) |
                           ((rw - bw * (px - s.ax)) < -s.r2));
        const bool in1 = (x < s.x1) & !(((ru - bu * (qx - s.bx)) < -s.r0) | ((rv - bv * (qx - s.cx)) < -s.r1) |
                                         ((rw - bw * (qx - s.ax)) < -s.r2));
        bits |= uint64_t(uint32_t(in0) | (uint32_t(in1) << 1)) << (8 * (y - s.y0) + (x - s.x0));
        x += 2;
        if (x > s.x1) {
            x = s.x0;
            ++y;
        }
    }
    return bits;
}

// Lane path for boxes at most 4 samples wide: one row of the box per step
// (a wave takes max(rows) steps). The x-dependent products of the four
// columns are computed once per face and the y-dependent ones once per row;
// every value is the reference expression's own rounding,
// u = (cx-bx)*(py-by) - (cy-by)*(px-bx), covered iff !(u+r0 < 0 || ...).
// Sample centres x + .5 are exact in double for every int x, so they are
// stepped by exact additions of 1. The row counter k is the same in every
// active lane, so each row's bits land at wave-uniform positions (bit
// 8*k + c for row k of [ya, yb]) and the bitmap is moved to the box's rows
// once at the end.
// Bitmap relative to (x0, y0), bit 8*(y-y0) + (x-x0), rows [ya, yb].
__device__ __forceinline__ uint64_t lane_bitmap4(const Setup& s, int ya, int yb) {
    const double au = s.cx - s.bx, bu = s.cy - s.by;
    const double av = s.ax - s.cx, bv = s.ay - s.cy;
    const double aw = s.bx - s.ax, bw = s.by - s.ay;
    const int w = s.x1 - s.x0 + 1;
    const double px0 = double(s.x0) + .5;
    double cu[4], cv[4], cw[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double px = px0 + double(c);
        cu[c] = bu * (px - s.bx);
        cv[c] = bv * (px - s.cx);
        cw[c] = bw * (px - s.ax);
    }
    // fl(u + r) < 0 tested as u < -r (see covers()).
    const double nu = -s.r0, nv = -s.r1, nw = -s.r2;
    const int rows = yb - ya + 1;
    if (rows <= 0) return 0;
    double py = double(ya) + .5;
    uint32_t lo = 0, hi = 0;
    for (int k = 0; k < rows; ++k) {
        const double ru = au * (py - s.by), rv = av * (py - s.cy), rw = aw * (py - s.ay);
        uint32_t row = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            // Bitwise | so that all three tests issue without branches.
            const bool in = (c < w) & !(((ru - cu[c]) < nu) | ((rv - cv[c]) < nv) | ((rw - cw[c]) < nw));
            row |= in ? 1u << c : 0u;
        }
        if (k < 4)
            lo |= row << (8 * k);
        else
            hi |= row << (8 * (k - 4));
        py += 1.0;
    }
    return ((uint64_t(hi) << 32) | lo) << (8 * (ya - s.y0));
}

// Largest v over the wave's lanes, for 0 <= v < 16, as a wave-uniform value
// (four ballots). Call with every lane of the wave active.
__device__ __forceinline__ int wave_max_small(int v) {
    int m = 0;
#pragma unroll
    for (int b = 8; b > 0; b >>= 1)
        if (__ballot(v >= m + b)) m += b;
    return m;
}

// lane_bitmap4 for a whole wave at once: C (3 or 4) columns and `max_rows`
// steps for every lane (both wave-uniform: the widest box and the most rows
// of the wave's lane-path faces), the lane's own width `w` and row count
// `rows` (0 for a lane with no such face) masking its bits. A uniform loop
// keeps the row counter and the bit positions in scalar registers, and a
// wave whose boxes are at most 3 wide (nearly all on fine meshes) skips the
// fourth column. Same samples, same rounding, same bitmap as lane_bitmap4.
template <int C>
__device__ __forceinline__ uint64_t lane_bitmap_wave(const Setup& s, int ya, int rows, int max_rows, int w) {
    const double au = s.cx - s.bx, bu = s.cy - s.by;
    const double av = s.ax - s.cx, bv = s.ay - s.cy;
    const double aw = s.bx - s.ax, bw = s.by - s.ay;
    const double px0 = double(s.x0) + .5;
    double cu[C], cv[C], cw[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const double px = px0 + double(c);
        cu[c] = bu * (px - s.bx);
        cv[c] = bv * (px - s.cx);
        cw[c] = bw * (px - s.ax);
    }
    // fl(u + r) < 0 tested as u < -r (see covers()).
    const double nu = -s.r0, nv = -s.r1, nw = -s.r2;
    double py = double(ya) + .5;
    uint32_t lo = 0, hi = 0;
    for (int k = 0; k < max_rows; ++k) {
        const double ru = au * (py - s.by), rv = av * (py - s.cy), rw = aw * (py - s.ay);
        const bool live = k < rows;
        uint32_t row = 0;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            // Bitwise | so that all three tests issue without branches.
            const bool in = live & (c < w) & !(((ru - cu[c]) < nu) | ((rv - cv[c]) < nv) | ((rw - cw[c]) < nw));
            row |= in ? 1u << c : 0u;
        }
        if (k < 4)
            lo |= row << (8 * k);
        else
            hi |= row << (8 * (k - 4));
        py += 1.0;
    }
    return ((uint64_t(hi) << 32) | lo) << (8 * (ya - s.y0));
}

// Samples [xa, xb] x [ya, yb] of one 4-column quad of a tile (columns q ..
// q + 3, q a multiple of 4; xa..xb within them; ya..yb within one tile row)
// covered by the face, by one lane: the products of the quad's four columns
// once, those of each row once per row, the reference expression's own
// rounding as in lane_bitmap4. Returns the tile word (bit 8*(y&7) + (x&7)).
__device__ __forceinline__ uint64_t quad_word(const Setup& s, int q, int xa, int xb, int ya, int yb) {
    const double au = s.cx - s.bx, bu = s.cy - s.by;
    const double av = s.ax - s.cx, bv = s.ay - s.cy;
    const double aw = s.bx - s.ax, bw = s.by - s.ay;
    const double px0 = double(q) + .5;
    double cu[4], cv[4], cw[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double px = px0 + double(c);
        cu[c] = bu * (px - s.bx);
        cv[c] = bv * (px - s.cx);
        cw[c] = bw * (px - s.ax);
    }
    // fl(u + r) < 0 tested as u < -r (see covers()).
    const double nu = -s.r0, nv = -s.r1, nw = -s.r2;
    const uint32_t columns = (0xFu >> (3 - (xb - q))) & (0xFu << (xa - q));
    const int shift = q & 7;
    uint64_t word = 0;
    double py = double(ya) + .5;
    for (int y = ya; y <= yb; ++y) {
        const double ru = au * (py - s.by), rv = av * (py - s.cy), rw = aw * (py - s.ay);
        uint32_t row = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            // Bitwise | so that all three tests issue without branches.
            const bool in = !(((ru - cu[c]) < nu) | ((rv - cv[c]) < nv) | ((rw - cw[c]) < nw));
            row |= in ? 1u << c : 0u;
        }
        word |= uint64_t((row & columns) << shift) << (8 * (y & 7));
        py += 1.0;
    }
    return word;
}

// One edge of a face along one sample row: the value at sample x is
// fl(t - fl(b * fl(x + .5 - p))) with t = fl(a * fl(py - q)) fixed for the
// row, and the sample passes the edge when !(value < nr), nr = -r -- the
// arithmetic of covers() and quad_word().
struct RowEdge {
    double t, b, p, nr;
};

__device__ __forceinline__ bool row_pass(const RowEdge& e, int x) {
    const double px = double(x) + .5;
    return !((e.t - e.b * (px - e.p)) < e.nr);
}

// Narrows [l, r] to its samples that pass edge e. fl(px - p), fl(b * d) and
// fl(t - m) are each monotone in their varying operand, so along a row the
// value never rises (b > 0) or never falls (b < 0) as x grows: the passing
// samples are a prefix (b > 0) or a suffix (b < 0) of [l, r], and all or
// none of it when b == 0. The end is guessed from the real crossing
// px = p + (t - nr) / b and then settled by testing samples, so the result
// is exact whatever the guess; a guess off by k samples costs k tests.
__device__ __forceinline__ void clip_row(const RowEdge& e, int& l, int& r) {
    if (l > r) return;
    if (!(e.b != 0)) {
        if (!row_pass(e, l)) r = l - 1;
        return;
    }
    // An approximate reciprocal: the guess only seeds the search.
    const double guess = e.p + (e.t - e.nr) * double(__builtin_amdgcn_rcpf(float(e.b))) - .5;
    if (e.b > 0) {
        // Last passing sample, l - 1 when none passes.
        int x = int(floor(fmin(fmax(guess, double(l - 1)), double(r))));
        if (x >= l && !row_pass(e, x)) {
            do --x;
            while (x >= l && !row_pass(e, x));
        } else {
            while (x < r && row_pass(e, x + 1)) ++x;
        }
        r = x;
    } else {
        // First passing sample, r + 1 when none passes.
        int x = int(ceil(fmin(fmax(guess, double(l)), double(r + 1))));
        if (x <= r && !row_pass(e, x)) {
            do ++x;
            while (x <= r && !row_pass(e, x));
        } else {
            while (x > l && row_pass(e, x - 1)) --x;
        }
        l = x;
    }
}

// In-place transpose of a 4x4 byte matrix held as four words: afterwards
// byte k of word j is what byte j of word k was (2-byte blocks swapped
// between words k and k+2, then bytes between words k and k+1).
__device__ __forceinline__ void transpose4(uint32_t& a0, uint32_t& a1, uint32_t& a2, uint32_t& a3) {
    uint32_t t = ((a0 >> 16) ^ a2) & 0x0000FFFFu;
    a2 ^= t;
    a0 ^= t << 16;
    t = ((a1 >> 16) ^ a3) & 0x0000FFFFu;
    a3 ^= t;
    a1 ^= t << 16;
    t = ((a0 >> 8) ^ a1) & 0x00FF00FFu;
    a1 ^= t;
    a0 ^= t << 8;
    t = ((a2 >> 8) ^ a3) & 0x00FF00FFu;
    a3 ^= t;
    a2 ^= t << 8;
}

// One half (four sample rows) of the eight tile words of one tile row
// under the samples [base, base + 63] (base a multiple of 64): the samples
// [xa, xb] x [ya, yb] of that span the face covers, the rows within one
// half of the tile row. w[c] gets bits 8 * (y & 3) + (x & 7) of tile
// base / 8 + c: the low half of its word for rows 0-3 of the tile, the high
// half for rows 4-7. With `full` (the rectangle classified as wholly
// covered) every row is [xa, xb]; otherwise each row's covered samples are
// the interval left by clip_row() over the three edges. Row r becomes a
// 64-bit column mask; its low four bytes are tiles 0-3 and its high four
// tiles 4-7, and two 4x4 byte transposes turn the rows into tile words.
__device__ __forceinline__ void span_half_words(const Setup& s, int base, int xa, int xb, int ya, int yb, bool full,
                                                uint32_t w[8]) {
    RowEdge eu{0, s.cy - s.by, s.bx, -s.r0};
    RowEdge ev{0, s.ay - s.cy, s.cx, -s.r1};
    RowEdge ew{0, s.by - s.ay, s.ax, -s.r2};
    const double au = s.cx - s.bx, av = s.ax - s.cx, aw = s.bx - s.ax;
    uint64_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
    double py = double(ya) + .5;
    for (int y = ya; y <= yb; ++y) {
        int l = xa, r = xb;
        if (!full) {
            eu.t = au * (py - s.by);
            ev.t = av * (py - s.cy);
            ew.t = aw * (py - s.ay);
            clip_row(eu, l, r);
            clip_row(ev, l, r);
            clip_row(ew, l, r);
        }
        const uint64_t mask = l <= r ? (~0ull >> (63 - (r - base))) & (~0ull << (l - base)) : 0;
        const int row = y & 3;
        m0 = row == 0 ? mask : m0;
        m1 = row == 1 ? mask : m1;
        m2 = row == 2 ? mask : m2;
        m3 = row == 3 ? mask : m3;
        py += 1.0;
    }
    w[0] = uint32_t(m0);
    w[1] = uint32_t(m1);
    w[2] = uint32_t(m2);
    w[3] = uint32_t(m3);
    w[4] = uint32_t(m0 >> 32);
    w[5] = uint32_t(m1 >> 32);
    w[6] = uint32_t(m2 >> 32);
    w[7] = uint32_t(m3 >> 32);
    transpose4(w[0], w[1], w[2], w[3]);
    transpose4(w[4], w[5], w[6], w[7]);
}

// Moves an 8x8 bitmap anchored at sample (x0, y0) into the up-to-four 8x8
// tile words it overlaps: word[0] tile (x0>>3, y0>>3), word[1] the tile to
// its right, word[2] below, word[3] below-right. Shifting the bitmap left by
// x0 & 7 puts each row's samples that stay in the left tile at columns
// >= x0 & 7 (`keep`, a byte mask repeated in every row); the others are the
// right tile's columns < x0 & 7, found by the complementary shift and mask.
__device__ __forceinline__ void scatter_bitmap(uint64_t local, int x0, int y0, uint64_t word[4]) {
    const uint32_t ox = uint32_t(x0) & 7u, oy = uint32_t(y0) & 7u;
    uint32_t keep = (0xFFu << ox) & 0xFFu;
    keep |= keep << 8;
    keep |= keep << 16;
    const uint64_t keep64 = (uint64_t(keep) << 32) | keep;
    const uint64_t left = (local << ox) & keep64;
    const uint64_t right = (local >> (8 - ox)) & ~keep64;
    word[0] = left << (8 * oy);
    word[1] = right << (8 * oy);
    word[2] = oy ? left >> (64 - 8 * oy) : 0;
    word[3] = oy ? right >> (64 - 8 * oy) : 0;
}

constexpr uint32_t kFar = 0xFFFFFFFFu;

// Horizontal distance from column x to the nearest site in one mask row,
// looking at most `reach` columns beyond x's own byte on each side; kFar if
// there is none that close. row_byte(t) is the row's 8 samples of tile column
// t (bit = x & 7).
template <class RowByte>
__device__ __forceinline__ uint32_t nearest_in_row_by(RowByte row_byte, uint32_t tiles, uint32_t x, uint32_t reach) {
    const uint32_t tx = x >> 3, bit = x & 7, base = tx << 3;
    const uint32_t byte = row_byte(tx);
    const uint32_t lower = byte & ((2u << bit) - 1), upper = byte & ~((1u << bit) - 1);
    uint32_t left = lower ? x - (base + 31 - __clz(lower)) : kFar;
    uint32_t right = upper ? (base + __ffs(upper) - 1) - x : kFar;
    if (left == kFar)
        for (int t = int(tx) - 1; t >= 0 && x - (uint32_t(t) * 8 + 7) <= reach; --t) {
            const uint32_t b = row_byte(uint32_t(t));
            if (b) {
                left = x - (uint32_t(t) * 8 + 31 - __clz(b));
                break;
            }
        }
    if (right == kFar)
        for (uint32_t t = tx + 1; t < tiles && t * 8 - x <= reach; ++t) {
            const uint32_t b = row_byte(t);
            if (b) {
                right = t * 8 + __ffs(b) - 1 - x;
                break;
            }
        }
    return min(left, right);
}

// nearest_in_row_by on row y of a tile-major mask.
__device__ __forceinline__ uint32_t nearest_in_row(const uint64_t* mask, uint32_t tiles, uint32_t x, uint32_t y,
                                                   uint32_t reach) {
    const uint64_t* row = mask + size_t(y >> 3) * tiles;
    const uint32_t shift = (y & 7) << 3;
    return nearest_in_row_by([&](uint32_t t) { return uint32_t(row[t] >> shift) & 0xFFu; }, tiles, x, reach);
}

// float(d*d) of the reference's row pass (src/evaluate.cpp:50), as an integer:
// exact below 2^24 and round-to-nearest-even above.
__device__ __forceinline__ uint32_t rounded_square(uint32_t d) {
    return uint32_t(float(d * d));
}

}  // namespace device
}  // namespace blitz_hip
