// Normals and attributes audit profiles on gfx950.
//
// Attribute raster (src/evaluate.cpp:125-189, attributed instantiation). The
// reference draws faces in index order and keeps, per sample, the last face
// whose double depth d passes `d < stored` where `stored` is the float the
// previous winner left (initially +inf). Let m be the smallest finite d over
// the faces strictly covering the sample and S = float(m). The stored value
// ends at S; the first face k with float(d) == S is always accepted, and after
// it exactly the faces with d < S are. So the winner is the largest face index
// with d < S if there is one, else k: a rule that any visiting order can
// apply incrementally (on a smaller S, restart k and that index). Here one
// wavefront owns one 8x8 tile (lane = sample), the faces binned to the tile
// are set up 64 at a time (one per lane, shared through LDS) and every lane
// applies the rule to each; the winner's barycentrics, normal, colour and
// material are then recomputed exactly as the reference computes them.
//
// Normal angle (src/evaluate.cpp:282-290): per view, the largest metric_acos
// of the clamped dot of the two rasters' normals over samples visible in
// both. Attributed distance (src/evaluate.cpp:217-270): per visible sample of
// one raster, the smallest product-metric cost to a visible sample of the
// other within the reference's window; the window's pruning only skips
// samples whose cost cannot be lower, so the minimum is the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "appearance_kernels.hpp"
#include "raster_device.hpp"

namespace blitz_hip {
namespace {

using namespace device;

constexpr uint32_t kNoFace = 0xFFFFFFFFu;

struct V3 {
    float x, y, z;
};

// include/blitz/mesh.hpp:76-81: dot in double, cross and add in float,
// scaling and normalisation through double.
__device__ __forceinline__ double dot3(V3 a, V3 b) {
    return double(a.x) * b.x + double(a.y) * b.y + double(a.z) * b.z;
}
__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 add3(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 scale3(V3 a, double s) { return {float(a.x * s), float(a.y * s), float(a.z * s)}; }
__device__ __forceinline__ V3 normalized3(V3 a) {
    const double l = sqrt(dot3(a, a));
    if (!(l > 0)) return {0, 0, 0};
    const double s = 1 / l;
    return {float(a.x * s), float(a.y * s), float(a.z * s)};
}
__device__ __forceinline__ double length3(V3 a) { return sqrt(dot3(a, a)); }
__device__ __forceinline__ V3 xyz(float4 v) { return {v.x, v.y, v.z}; }

// fdlibm e_acos.c notice, preserved as its licence requires:
// Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
// Developed at SunSoft, a Sun Microsystems, Inc. business.
// Permission to use, copy, modify, and distribute this software is freely
// granted, provided that this notice is preserved.
// The routine: a rational
// approximation of asin about 0, half-angle forms near +-1 with the split
// square root for x > 0.5, in the operation order of src/metric_angle.hpp:21-40
// so that every rounding is the reference's.
__device__ double metric_acos(double x) {
    const double pi = 3.14159265358979311600e+00, half = 1.57079632679489655800e+00,
                 lo = 6.12323399573676603587e-17;
    if (x == 1) return 0;
    if (x == -1) return pi + 2 * lo;
    const double z = x < -.5 ? (1 + x) * .5 : x > .5 ? (1 - x) * .5 : x * x;
    const double p = z * (1.66666666666666657415e-01 +
                          z * (-3.25565818622400915405e-01 +
                               z * (2.01212532134862925881e-01 +
                                    z * (-4.00555345006794114027e-02 +
                                         z * (7.91534994289814532176e-04 + z * 3.47933107596021167570e-05)))));
    const double q = 1 + z * (-2.40339491173441421878e+00 +
                              z * (2.02094576023350569471e+00 +
                                   z * (-6.88283971605453293030e-01 + z * 7.70381505559019352791e-02)));
    const double r = p / q;
    if (x >= -.5 && x <= .5) {
        if (x > -0x1p-57 && x < 0x1p-57) return half + lo;
        return half - (x - (lo - x * r));
    }
    const double s = sqrt(z);
    if (x < 0) return pi - 2 * (s + (r * s - lo));
    const double high = __longlong_as_double(__double_as_longlong(s) & (long long)0xffffffff00000000ull);
    const double correction = (z - high * high) / (s + high);
    return 2 * (high + (r * s + correction));
}

// std::clamp(v, -1.0, 1.0) and std::max(1e-30, v) as the reference's
// library defines them (a NaN passes clamp unchanged and loses to 1e-30).
__device__ __forceinline__ double clamp1(double v) { return v < -1.0 ? -1.0 : (1.0 < v ? 1.0 : v); }
__device__ __forceinline__ double max_tiny(double v) { return 1e-30 < v ? v : 1e-30; }

__device__ __forceinline__ uint32_t tile_sample(uint32_t tiles, uint32_t x, uint32_t y) {
    return ((y >> 3) * tiles + (x >> 3)) * 64 + (y & 7) * 8 + (x & 7);
}

// One binned face after setup: corners after the back-face swap, their view
// depths, the positive area, the sample box and the top-left flags of the
// edges bc, ca, ab (bits 0, 1, 2).
struct FaceRec {
    double ax, ay, bx, by, cx, cy;
    double za, zb, zc, area;
    int x0, x1, y0, y1;
    uint32_t face;
    uint32_t top_left;
};

constexpr int kWaves = 4;

// Double depth of face r at sample centre (px, py) when the sample is
// strictly inside it (src/evaluate.cpp:171-181); false otherwise.
__device__ __forceinline__ bool inside_depth(const FaceRec& r, double px, double py, bool persp, double diameter,
                                             double& depth) {
    double u = edge(r.bx, r.by, r.cx, r.cy, px, py);
    double v = edge(r.cx, r.cy, r.ax, r.ay, px, py);
    double w = edge(r.ax, r.ay, r.bx, r.by, px, py);
    if (u < 0 || v < 0 || w < 0) return false;
    u /= r.area;
    v /= r.area;
    w /= r.area;
    if (persp) {
        const double sum = u / r.za + v / r.zb + w / r.zc;
        depth = 1 / sum;
    } else {
        depth = u * r.za + v * r.zb + w * r.zc;
    }
    depth /= diameter;
    return true;
}

// Setup of binned face f (src/evaluate.cpp:144-165) as a FaceRec, with the
// top-left edge test of src/neural/evaluate.cu:198-200; depths are loaded when
// zs is given.
__device__ __forceinline__ bool top_left(double ax, double ay, double bx, double by) {
    return by < ay || (by == ay && bx < ax);
}
__device__ FaceRec face_record(const AttrTask& T, uint32_t f, global_ptr<double> zs) {
    const RasterTask& t = T.raster;
    const uint3 tri = load_face(t, f);
    const Corners c = load_corners(t, tri, f);
    Setup s;
    FaceRec r;
    r.face = f;
    if (setup_triangle(t, c, s, double(t.size)) == 1) {
        const double area = edge(c.a.x, c.a.y, c.b.x, c.b.y, c.c.x, c.c.y);
        const bool back = area < 0;
        r.ax = s.ax;
        r.ay = s.ay;
        r.bx = s.bx;
        r.by = s.by;
        r.cx = s.cx;
        r.cy = s.cy;
        r.area = back ? -area : area;
        if (zs) {
            r.za = zs[tri.x];
            r.zb = back ? zs[tri.z] : zs[tri.y];
            r.zc = back ? zs[tri.y] : zs[tri.z];
        }
        r.x0 = s.x0;
        r.x1 = s.x1;
        r.y0 = s.y0;
        r.y1 = s.y1;
        r.top_left = uint32_t(top_left(s.bx, s.by, s.cx, s.cy)) | uint32_t(top_left(s.cx, s.cy, s.ax, s.ay)) << 1 |
                uint32_t(top_left(s.ax, s.ay, s.bx, s.by)) << 2;
    } else {  // binned faces are drawable; keep an empty box regardless
        r.x0 = r.y0 = 1;
        r.x1 = r.y1 = 0;
    }
    return r;
}

__global__ __launch_bounds__(256) void attribute_kernel(const AttrTask* __restrict__ tasks) {
    __shared__ FaceRec recs[kWaves][64];
    const AttrTask& T = tasks[blockIdx.y];
    const RasterTask& t = T.raster;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t words = t.tiles * t.tiles;
    const uint32_t tile = blockIdx.x * kWaves + wave;
    if (tile >= words) return;  // whole waves exit together
    const uint32_t tx = tile % t.tiles, ty = tile / t.tiles;
    const int x = int(tx * 8 + (lane & 7)), y = int(ty * 8 + (lane >> 3));
    const bool real = x < int(t.size) && y < int(t.size);
    const double px = x + .5, py = y + .5;
    const bool persp = T.perspective != 0;
    const unsigned long long b0 = T.bin_offsets[tile], b1 = T.bin_offsets[tile + 1];
    const global_ptr<double> zs = global(T.depth_z);
    float S = INFINITY;
    uint32_t k = kNoFace, last = 0;  // last = 1 + largest face with d < S, 0 if none
    FaceRec* mine = recs[wave];
    for (unsigned long long base = b0; base < b1; base += 64) {
        const uint32_t n = uint32_t(min(64ull, b1 - base));
        if (lane < n) mine[lane] = face_record(T, T.bin_faces[base + lane], zs);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (real)
            for (uint32_t j = 0; j < n; ++j) {
                const FaceRec& r = mine[j];
                if (x < r.x0 || x > r.x1 || y < r.y0 || y > r.y1) continue;
                double d;
                if (!inside_depth(r, px, py, persp, T.diameter, d)) continue;
                if (!(d < double(INFINITY))) continue;  // never passes `d < stored`
                const float fd = float(d);
                if (fd < S) {
                    S = fd;
                    k = r.face;
                    last = d < double(fd) ? r.face + 1 : 0;
                } else if (fd == S) {
                    k = min(k, r.face);
                    if (d < double(S)) last = max(last, r.face + 1);
                }
            }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    const uint32_t out = tile * 64 + lane;
    const uint32_t win = last ? last - 1 : k;
    if (!real || win == kNoFace) {
        T.normal_out[out] = make_float4(0, 0, 0, 0);
        if (T.color_out) T.color_out[out] = make_float4(0, 0, 0, __uint_as_float(0));
        if (T.depth_out) T.depth_out[out] = INFINITY;
        return;
    }
    // The winner, exactly as the reference writes it (src/evaluate.cpp:144-184).
    const uint3 tri = load_face(t, win);
    const Corners c = load_corners(t, tri, win);
    double area = edge(c.a.x, c.a.y, c.b.x, c.b.y, c.c.x, c.c.y);
    const bool back = area < 0;
    const uint32_t id = back ? tri.z : tri.y, ie = back ? tri.y : tri.z;
    const double2 A = c.a, D = back ? c.c : c.b, E = back ? c.b : c.c;
    if (back) area = -area;
    double u = edge(D.x, D.y, E.x, E.y, px, py), v = edge(E.x, E.y, A.x, A.y, px, py),
           w = edge(A.x, A.y, D.x, D.y, px, py);
    u /= area;
    v /= area;
    w /= area;
    if (persp) {
        const double za = zs[tri.x], zd = zs[id], ze = zs[ie];
        const double sum = u / za + v / zd + w / ze;
        u = u / za / sum;
        v = v / zd / sum;
        w = w / ze / sum;
    }
    // Geometric normal from the unswapped corners, scaled by the diameter.
    const float4 p0 = T.positions[tri.x], p1 = T.positions[tri.y], p2 = T.positions[tri.z];
    const double diameter = T.diameter;
    const V3 d1{float((double(p1.x) - p0.x) / diameter), float((double(p1.y) - p0.y) / diameter),
                float((double(p1.z) - p0.z) / diameter)};
    const V3 d2{float((double(p2.x) - p0.x) / diameter), float((double(p2.y) - p0.y) / diameter),
                float((double(p2.z) - p0.z) / diameter)};
    V3 geometric = normalized3(cross3(d1, d2));
    if (back) geometric = scale3(geometric, -1);
    V3 na{0, 0, 0}, nd{0, 0, 0}, ne{0, 0, 0};
    if (T.normals) {
        na = xyz(T.normals[tri.x]);
        nd = xyz(T.normals[id]);
        ne = xyz(T.normals[ie]);
    }
    V3 n = normalized3(add3(add3(scale3(na, u), scale3(nd, v)), scale3(ne, w)));
    if (length3(n) < .5)
        n = geometric;
    else if (back)
        n = scale3(n, -1);
    float4 col = make_float4(1, 1, 1, 0);
    if (T.colors) {
        const float4 ca = T.colors[tri.x], cd = T.colors[id], ce = T.colors[ie];
        col.x = float(u * ca.x + v * cd.x + w * ce.x);
        col.y = float(u * ca.y + v * cd.y + w * ce.y);
        col.z = float(u * ca.z + v * cd.z + w * ce.z);
    }
    const uint32_t material = T.materials ? T.materials[win] : 0u;
    col.w = __uint_as_float(material);
    T.normal_out[out] = make_float4(n.x, n.y, n.z, 1.f);
    if (T.color_out) T.color_out[out] = col;
    if (T.depth_out) T.depth_out[out] = S;
}

// Overlap (src/neural/evaluate.cu:301-309,352-355): per sample, the faces
// whose interior holds its centre under the top-left fill rule; per view, the
// sum of those counts and the number of samples with a non-zero count.
__global__ __launch_bounds__(256) void overlap_kernel(const AttrTask* __restrict__ tasks,
                                                      unsigned long long* __restrict__ sums) {
    __shared__ FaceRec recs[kWaves][64];
    const AttrTask& T = tasks[blockIdx.y];
    const RasterTask& t = T.raster;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t words = t.tiles * t.tiles;
    const uint32_t tile = blockIdx.x * kWaves + wave;
    if (tile >= words) return;  // whole waves exit together
    const uint32_t tx = tile % t.tiles, ty = tile / t.tiles;
    const int x = int(tx * 8 + (lane & 7)), y = int(ty * 8 + (lane >> 3));
    const bool real = x < int(t.size) && y < int(t.size);
    const double px = x + .5, py = y + .5;
    const unsigned long long b0 = T.bin_offsets[tile], b1 = T.bin_offsets[tile + 1];
    uint32_t count = 0;
    FaceRec* mine = recs[wave];
    for (unsigned long long base = b0; base < b1; base += 64) {
        const uint32_t n = uint32_t(min(64ull, b1 - base));
        if (lane < n) mine[lane] = face_record(T, T.bin_faces[base + lane], global_ptr<double>(nullptr));
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (real)
            for (uint32_t j = 0; j < n; ++j) {
                const FaceRec& r = mine[j];
                if (x < r.x0 || x > r.x1 || y < r.y0 || y > r.y1) continue;
                const double u = edge(r.bx, r.by, r.cx, r.cy, px, py);
                const double v = edge(r.cx, r.cy, r.ax, r.ay, px, py);
                const double w = edge(r.ax, r.ay, r.bx, r.by, px, py);
                if (u < 0 || v < 0 || w < 0) continue;
                if ((u > 0 || (r.top_left & 1)) && (v > 0 || (r.top_left & 2)) && (w > 0 || (r.top_left & 4))) ++count;
            }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    unsigned long long samples = count, pixels = count != 0;
    for (int o = 32; o > 0; o >>= 1) {
        samples += __shfl_xor(samples, o);
        pixels += __shfl_xor(pixels, o);
    }
    if (lane == 0 && pixels) {
        atomicAdd(&sums[2 * blockIdx.y], samples);
        atomicAdd(&sums[2 * blockIdx.y + 1], pixels);
    }
}

__device__ __forceinline__ double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// Largest value of a 256-thread block, valid in thread 0.
__device__ __forceinline__ double block_max(double v) {
    __shared__ double part[4];
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    return fmax(fmax(part[0], part[1]), fmax(part[2], part[3]));
}

__device__ __forceinline__ double load_max(const unsigned long long* p) {
    return __longlong_as_double((long long)__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

constexpr uint32_t kReduceBlocks = 2048;  // blocks per raster pair of the streaming reductions

// Grid (blocks, pairs); each block strides over the pair's samples.
__global__ __launch_bounds__(256) void normal_angle_kernel(const AttrPair* __restrict__ pairs) {
    const AttrPair& p = pairs[blockIdx.y];
    const uint32_t n = p.tiles * p.tiles * 64;
    double a = 0;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float4 x = p.normal_a[i], y = p.normal_c[i];
        if (x.w != 0 && y.w != 0) {
            const double v = metric_acos(clamp1(dot3(xyz(x), xyz(y))));
            if (v > a) a = v;  // a NaN never raises the maximum
        }
    }
    a = block_max(a);
    if (threadIdx.x == 0 && a > 0) atomicMax(&p.out[0], (unsigned long long)__double_as_longlong(a));
}

// src/evaluate.cpp:83-95
__device__ __forceinline__ double sample_cost(float4 na, float4 ca, float4 nb, float4 cb, double spatial,
                                              const CostSettings& s) {
    double cost = spatial;
    if (s.profile != 0 && s.normal_weight > 0) {
        const V3 a = xyz(na), b = xyz(nb);
        const double cosine = clamp1(dot3(a, b) / max_tiny(length3(a) * length3(b)));
        const double angle = metric_acos(cosine) * s.normal_weight;
        cost += angle * angle;
    }
    if (s.profile == 2) {
        const double x = double(ca.x) - cb.x, y = double(ca.y) - cb.y, z = double(ca.z) - cb.z;
        cost += s.color_weight * s.color_weight * (x * x + y * y + z * z);
        if (__float_as_uint(ca.w) != __float_as_uint(cb.w)) cost += s.material_weight * s.material_weight;
    }
    return cost;
}

// Grid (blocks, pairs, 2 directions), launched twice (phase 0: samples whose
// own position is not visible in the other raster; phase 1: the others);
// each thread strides over samples of the source raster and, for each
// visible one, finds the reference's windowed minimum cost
// (src/evaluate.cpp:229-262). out[1] holds the largest
// minimum found so far, shared by both directions as the reference's
// `maximum` is. A sample whose running minimum is already <= that value
// cannot raise it and stops (the early break of Taha and Hanbury, "An
// efficient algorithm for calculating the exact Hausdorff distance", IEEE
// TPAMI 2015): out[1] ends at the reference's maximum whenever that is within
// limit^2, and above limit^2 (an infinite distance) otherwise.
__global__ __launch_bounds__(256) void attributed_kernel(const AttrPair* __restrict__ pairs, CostSettings s,
                                                         uint32_t phase) {
    const AttrPair& p = pairs[blockIdx.y];
    const bool reverse = blockIdx.z != 0;
    const float4* fn = reverse ? p.normal_c : p.normal_a;
    const float4* fc = reverse ? p.color_c : p.color_a;
    const float4* tn = reverse ? p.normal_a : p.normal_c;
    const float4* tc = reverse ? p.color_a : p.color_c;
    const uint32_t n = p.tiles * p.tiles * 64;
    const int size = int(p.size);
    const double limit2 = s.limit * s.limit, invs2 = 1.0 / double(s.ss * s.ss);
    const int radius = int(fmin(double(size), ceil(s.limit * s.ss)));
    unsigned long long* slot = &p.out[1];
    // Colours and materials enter the cost under the attributes profile only.
    const bool colored = s.profile == 2;
    const float4 none = make_float4(0, 0, 0, 0);
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        if (fn[i].w == 0) continue;
        // Phase 0 takes the samples whose own position is not visible in the
        // other raster (the likeliest largest minima), phase 1 the rest.
        const bool seen = tn[i].w != 0;
        if (seen != (phase != 0)) continue;
        const float4 an = fn[i], ac = colored ? fc[i] : none;
        double best = INFINITY;
        if (seen) best = sample_cost(an, ac, tn[i], colored ? tc[i] : none, 0, s);
        if (best <= 1e-18) continue;
        double known = load_max(slot);
        if (known > limit2) return;  // the distance is already infinite
        if (best <= known) continue;
        const uint32_t tile = i >> 6, lane = i & 63;
        const int x = int((tile % p.tiles) * 8 + (lane & 7)), y = int((tile / p.tiles) * 8 + (lane >> 3));
        const int r = int(fmin(double(radius), ceil(sqrt(fmin(best, limit2)) * s.ss)));
        // The reference scans the (2r+1)^2 window row by row; its minimum
        // does not depend on the order, so the window is walked in square
        // rings around the sample. Every sample of ring k has spatial >=
        // k*k*invs2 (the same rounded expression at dx = k, dy = 0), and the
        // reference skips samples with spatial >= best or > limit^2, so the
        // walk ends at the first ring past either.
        auto visit = [&](int xx, int yy) {
            if (xx < 0 || yy < 0 || xx >= size || yy >= size) return;
            const uint32_t j = tile_sample(p.tiles, uint32_t(xx), uint32_t(yy));
            const float4 bn = tn[j];
            if (bn.w == 0) return;
            const double dx = x - xx, dy = y - yy, spatial = (dx * dx + dy * dy) * invs2;
            if (spatial >= best || spatial > limit2) return;
            const double cost = sample_cost(an, ac, bn, colored ? tc[j] : none, spatial, s);
            if (cost < best) best = cost;
        };
        for (int k = 1; k <= r; ++k) {
            const double ring = double(k * k) * invs2;
            if (ring >= best || ring > limit2) break;
            for (int d = -k; d <= k; ++d) {
                visit(x + d, y - k);
                visit(x + d, y + k);
            }
            for (int d = -k + 1; d <= k - 1; ++d) {
                visit(x - k, y + d);
                visit(x + k, y + d);
            }
            known = load_max(slot);
            if (best <= known) break;
        }
        if (best > known) atomicMax(slot, (unsigned long long)__double_as_longlong(best));
    }
}

}  // namespace

void launch_attribute_raster(const AttrTask* tasks, uint32_t count, uint32_t tiles, hipStream_t stream) {
    if (!count) return;
    const uint32_t blocks = (tiles * tiles + kWaves - 1) / kWaves;
    attribute_kernel<<<dim3(blocks, count), dim3(64 * kWaves), 0, stream>>>(tasks);
}

// About eight samples per thread, at most 2048 blocks per raster pair, so a
// few views still fill the chip.
uint32_t reduce_blocks(uint32_t tiles) { return std::min(kReduceBlocks, (tiles * tiles * 64 + 2047) / 2048); }

void launch_overlap(const AttrTask* tasks, uint32_t count, uint32_t tiles, unsigned long long* sums,
                    hipStream_t stream) {
    if (!count) return;
    const uint32_t blocks = (tiles * tiles + kWaves - 1) / kWaves;
    overlap_kernel<<<dim3(blocks, count), dim3(64 * kWaves), 0, stream>>>(tasks, sums);
}

void launch_normal_angle(const AttrPair* pairs, uint32_t count, uint32_t tiles, hipStream_t stream) {
    if (!count) return;
    normal_angle_kernel<<<dim3(reduce_blocks(tiles), count), dim3(256), 0, stream>>>(pairs);
}

void launch_attributed(const AttrPair* pairs, uint32_t count, uint32_t tiles, const CostSettings& s,
                       hipStream_t stream) {
    if (!count) return;
    // One thread per 4 samples: the windowed searches are latency-bound
    // gathers, so the launch keeps many waves resident.
    const uint32_t blocks = std::min<uint32_t>(16384, (tiles * tiles * 64 + 1023) / 1024);
    for (uint32_t phase = 0; phase < 2; ++phase)
        attributed_kernel<<<dim3(blocks, count, 2), dim3(256), 0, stream>>>(pairs, s, phase);
}

}  // namespace blitz_hip
