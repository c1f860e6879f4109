// CPU oracle for the coverage visual audit.
//
// TEST INFRASTRUCTURE, NOT PRODUCT CODE. Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg load this library, and only as the checker or
// as the timed CPU baseline. The HIP product path never calls into it.
//
// It restates, in plain single-threaded C++, the reference CPU audit for the
// coverage, normals and attributes profiles (reference: src/evaluate.cpp,
// src/mesh.cpp, src/evaluation_settings.hpp, src/metric_angle.hpp). Every function cites the reference lines it
// follows. Arithmetic is FP64 in the same expression order as the reference and
// is compiled with -ffp-contract=off, so results are bit-identical to a build of
// the reference sources (pinned by tests/test_oracle.py against the fixtures in
// tests/golden/, which come from that build).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <type_traits>
#include <vector>

// ORACLE_PLANT selects a deliberately wrong variant, built only by
// tests/test_oracle.py to show that the pins catch plausible mistakes:
// 1 drops the conservative radius of edge 0, 2 transposes an edge function's
// operands, 3 keeps the row pass in double (one float rounding instead of the
// reference's two), 4 breaks worst-view ties towards the later view,
// 5 leaves back-face normals unflipped, 6 drops the material penalty, 7
// counts pixel centres on every edge in the overlap count (no top-left rule).
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif

namespace {
constexpr double kPi = 3.14159265358979323846;
constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr uint64_t kMaxSamples = 64000000;  // src/evaluate.cpp:129

struct F3 { float x, y, z; };

// Layout mirrors blitz::Camera (include/blitz/evaluate.hpp:46).
struct Camera {
    F3 right, up, forward;
    double distance, focal, scale;
    uint8_t perspective;
};

// Layout mirrors blitz::Measurement (include/blitz/evaluate.hpp:34-39) with
// bools as bytes.
struct Measurement {
    double error, coverage, coverage_upper, changed_area, normal_degrees;
    uint32_t worst_view, changed_area_worst_view, views_evaluated;
    uint8_t supersample, complete, passed, resource_limited, cancelled;
};

struct Settings {
    double screen_size, limit, max_changed_area;
    uint32_t rotation_seed;
    uint16_t orthographic, perspective;
    uint8_t supersample, max_supersample, force_two_sided, conservative_screen;
    int32_t profile;
    double normal_weight, color_weight, material_weight;
};

struct Mesh {
    const float* positions;  // xyz triples, tightly packed
    uint64_t vertices;
    const uint32_t* indices;
    uint64_t index_count;
    const uint16_t* materials;  // optional, one per face
    uint64_t material_count;
    const uint8_t* double_sided;  // optional, indexed by material
    uint64_t double_sided_count;
    const float* normals = nullptr;         // optional xyz per vertex
    const uint8_t* colors = nullptr;        // optional RGBA8 per vertex
    F3 vertex(uint64_t i) const { return {positions[3 * i], positions[3 * i + 1], positions[3 * i + 2]}; }
    uint16_t material(uint64_t f) const { return material_count ? materials[f] : 0; }
    uint64_t faces() const { return index_count / 3; }
    bool two_sided(uint64_t f) const {  // include/blitz/mesh.hpp:103-104
        uint16_t m = material_count ? materials[f] : 0;
        return m < double_sided_count && double_sided[m];
    }
};

struct Bounds { F3 center; double radius; };

// include/blitz/mesh.hpp:78
double dot(F3 a, F3 b) { return double(a.x) * b.x + double(a.y) * b.y + double(a.z) * b.z; }
// include/blitz/mesh.hpp:79 (float arithmetic)
F3 cross(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// include/blitz/mesh.hpp:77,80-81
F3 normalized(F3 a) {
    double l = std::sqrt(dot(a, a));
    if (!(l > 0)) return {0, 0, 0};
    double s = 1 / l;
    return {float(a.x * s), float(a.y * s), float(a.z * s)};
}

// src/evaluate.cpp:16-20 — seed-dependent rotation of the Fibonacci directions.
F3 rotate(F3 v, uint32_t seed) {
    const double a = (seed % 65521) * .000731, b = ((seed >> 8) % 32749) * .000913;
    const double x = v.x * std::cos(a) - v.y * std::sin(a);
    const double y = v.x * std::sin(a) + v.y * std::cos(a);
    return {float(x), float(y * std::cos(b) - v.z * std::sin(b)), float(y * std::sin(b) + v.z * std::cos(b))};
}

// src/evaluate.cpp:21
inline double edge(double ax, double ay, double bx, double by, double x, double y) {
    return (bx - ax) * (y - ay) - (by - ay) * (x - ax);
}

// src/evaluate.cpp:108-124
std::vector<Camera> make_cameras(const Bounds& b, double screen, uint16_t ortho, uint16_t persp, uint32_t seed) {
    std::vector<Camera> out;
    for (int projection = 0; projection < 2; ++projection) {
        const unsigned count = projection ? persp : ortho;
        for (unsigned i = 0; i < count; ++i) {
            const double z = 1 - 2 * (i + .5) / count;
            const double r = std::sqrt(std::max(0.0, 1 - z * z));
            const double angle = i * kPi * (3 - std::sqrt(5.0));
            const F3 dir{float(r * std::cos(angle)), float(z), float(r * std::sin(angle))};
            const F3 forward = normalized(rotate(dir, seed + uint32_t(projection) * 104729u));
            const F3 hint = std::abs(forward.y) > .99 ?F3{1, 0, 0} : F3{0, 1, 0};
            const F3 right = normalized(cross(hint, forward));
            const F3 up = normalized(cross(forward, right));
            const double height = std::max(1024.0, 2 * screen);
            const double focal = height / (2 * std::tan(kPi / 6));
            const double distance = b.radius * std::sqrt(1 + std::pow(2 * focal / screen, 2));
            out.push_back({right, up, forward, distance, focal, screen / (b.radius * 2), uint8_t(projection)});
        }
    }
    return out;
}

struct Mask {
    uint32_t size = 0;
    std::vector<uint64_t> words;  // row-major bits, index = y*size + x (src/coverage.hpp:6-11)
    bool clipped = false;
    bool bit(size_t i) const { return (words[i / 64] >> (i % 64)) & 1; }
};

struct LengthError {};

// src/evaluate.cpp:125-189, coverage-only instantiation (CoverageRaster).
Mask rasterize(const Mesh& m, const Bounds& b, const Camera& c, double screen, uint8_t ss, bool force_two) {
    if (!ss || !std::isfinite(screen) || screen <= 0 || screen > 16384) throw std::invalid_argument("invalid raster extent");
    const uint32_t size = uint32_t(std::ceil(screen + 8)) * ss;
    if (uint64_t(size) * size > kMaxSamples) throw LengthError{};
    Mask out;
    out.size = size;
    out.words.assign((size_t(size) * size + 63) / 64, 0);
    struct P { double x, y, z; };
    std::vector<P> proj(m.vertices);
    for (uint64_t i = 0; i < m.vertices; ++i) {
        const F3 a = m.vertex(i);
        const double x = double(a.x) - b.center.x, y = double(a.y) - b.center.y, z0 = double(a.z) - b.center.z;
        auto along = [&](F3 axis) { return x * axis.x + y * axis.y + z0 * axis.z; };
        const double z = c.distance - along(c.forward);
        const double scale = c.perspective ? c.focal / z : c.scale;
        proj[i] = {along(c.right) * scale * ss + size * .5, along(c.up) * scale * ss + size * .5, z};
    }
    for (uint64_t f = 0; f < m.faces(); ++f) {
        const uint32_t ia = m.indices[3 * f];
        P a = proj[ia], d = proj[m.indices[3 * f + 1]], e = proj[m.indices[3 * f + 2]];
        if (c.perspective && (a.z <= 0 || d.z <= 0 || e.z <= 0)) { out.clipped = true; continue; }
        double area = edge(a.x, a.y, d.x, d.y, e.x, e.y);
        if (std::abs(area) < 1e-16) continue;
        const bool back = area < 0, two = force_two || m.two_sided(f);
        if (back && !two) continue;
        if (back) { std::swap(d, e); area = -area; }
        const double minx = std::min({a.x, d.x, e.x}), maxx = std::max({a.x, d.x, e.x});
        const double miny = std::min({a.y, d.y, e.y}), maxy = std::max({a.y, d.y, e.y});
        if (!std::isfinite(minx) || !std::isfinite(maxx) || !std::isfinite(miny) || !std::isfinite(maxy) ||
            minx < 0 || miny < 0 || maxx > size || maxy > size) {
            out.clipped = true;
            return out;  // the reference stops at the first out-of-frame face
        }
        const int x0 = std::max(0, int(std::floor(minx))), x1 = std::min(int(size) - 1, int(std::floor(maxx)));
        const int y0 = std::max(0, int(std::floor(miny))), y1 = std::min(int(size) - 1, int(std::floor(maxy)));
        const double r0 = ORACLE_PLANT == 1 ? 0.0 : .5 * (std::abs(e.x - d.x) + std::abs(e.y - d.y));
        const double r1 = .5 * (std::abs(a.x - e.x) + std::abs(a.y - e.y));
        const double r2 = .5 * (std::abs(d.x - a.x) + std::abs(d.y - a.y));
        for (int y = y0; y <= y1; ++y)
            for (int x = x0; x <= x1; ++x) {
                const double u = edge(d.x, d.y, e.x, e.y, x + .5, y + .5);
                const double v = ORACLE_PLANT == 2 ? edge(e.y, e.x, a.y, a.x, y + .5, x + .5)
                                                   : edge(e.x, e.y, a.x, a.y, x + .5, y + .5);
                const double w = edge(a.x, a.y, d.x, d.y, x + .5, y + .5);
                if (u + r0 < 0 || v + r1 < 0 || w + r2 < 0) continue;
                const size_t index = size_t(y) * size + x;
                out.words[index / 64] |= uint64_t(1) << (index % 64);
            }
    }
    return out;
}

// Felzenszwalb & Huttenlocher (2012), "Distance Transforms of Sampled
// Functions", 1-D lower envelope of parabolas; src/evaluate.cpp:33-53 uses the
// same FP64 intersection formula and writes float(d*d + f).
template <class In, class Out>
void envelope_1d(const In* f, Out* out, int n, std::vector<int>& site, std::vector<double>& bound) {
    int k = 0;
    site[0] = 0;
    bound[0] = -kInf;
    bound[1] = kInf;
    for (int q = 1; q < n; ++q) {
        double s;
        while (true) {
            const int p = site[k];
            s = ((double(f[q]) + double(q) * q) - (double(f[p]) + double(p) * p)) / (2 * (q - p));
            if (s > bound[k] || k == 0) break;
            --k;
        }
        ++k;
        site[k] = q;
        bound[k] = s;
        bound[k + 1] = kInf;
    }
    k = 0;
    for (int q = 0; q < n; ++q) {
        while (bound[k + 1] < q) ++k;
        const double d = q - site[k];
        out[q] = Out(d * d + f[site[k]]);
    }
}

// src/evaluate.cpp:58-70: separable squared EDT; sites are covered samples
// (value 0), everything else starts at 1e15f.
std::vector<float> distance_field(const Mask& r) {
    const size_t w = r.size, h = r.size, n = w * h;
    using Row = std::conditional_t<ORACLE_PLANT == 3, double, float>;
    std::vector<float> a(n);
    std::vector<Row> b(n);
    for (size_t i = 0; i < n; ++i) a[i] = r.bit(i) ? 0.f : 1e15f;
    const size_t len = std::max(w, h);
    std::vector<int> site(len);
    std::vector<double> bound(len + 1);
    std::vector<Row> col(len);
    std::vector<float> res(len);
    for (size_t y = 0; y < h; ++y) envelope_1d(a.data() + y * w, b.data() + y * w, int(w), site, bound);
    for (size_t x = 0; x < w; ++x) {
        for (size_t y = 0; y < h; ++y) col[y] = b[y * w + x];
        envelope_1d(col.data(), res.data(), int(h), site, bound);
        for (size_t y = 0; y < h; ++y) a[y * w + x] = res[y];
    }
    return a;
}

// src/evaluate.cpp:196-209 (coverage_distance_with) for packed masks.
double coverage_distance(const Mask& a, const Mask& b, uint8_t ss) {
    bool ca = false, cb = false, same = true;
    for (size_t i = 0; i < a.words.size(); ++i) {
        ca |= a.words[i] != 0;
        cb |= b.words[i] != 0;
        same &= a.words[i] == b.words[i];
    }
    if (same) return 0;
    if (ca != cb) return kInf;
    double maximum = 0;
    const size_t n = size_t(a.size) * a.size;
    auto directed = [&](const Mask& from, const Mask& to) {
        const std::vector<float> field = distance_field(to);
        for (size_t i = 0; i < n; ++i)
            if (from.bit(i)) maximum = std::max(maximum, double(field[i]));
    };
    directed(a, b);
    directed(b, a);
    return std::sqrt(maximum) / ss;
}

// src/evaluate.cpp:271-298, coverage profile.
void finish_view(Measurement& cur, const Mask& a, const Mask& c, const Settings& s, uint8_t ss, double coverage) {
    cur.supersample = ss;
    cur.coverage = coverage;
    cur.coverage_upper = cur.coverage + 2 * std::sqrt(2.0) / ss + 1e-6;
    if (c.clipped || a.clipped) cur.coverage_upper = kInf;
    uint64_t changed = 0, total = 0;
    for (size_t i = 0; i < a.words.size(); ++i) {
        changed += __builtin_popcountll(a.words[i] ^ c.words[i]);
        total += __builtin_popcountll(a.words[i] | c.words[i]);
    }
    cur.error = cur.coverage_upper;
    cur.changed_area = total ? double(changed) / total : 0;
    cur.passed = cur.error <= s.limit && cur.changed_area <= s.max_changed_area;
    if (s.conservative_screen) {
        cur.error = (a.clipped || c.clipped) ? 0 : std::max(0.0, cur.coverage - 2 * std::sqrt(2.0) / ss - 1e-6);
        cur.passed = cur.error <= s.limit;
    }
}

// ---------------------------------------------------------------- appearance

struct F4 { float x, y, z, w; };

// include/blitz/evaluate.hpp:40-44
struct Pixel {
    F3 normal{0, 0, 0};
    F4 color{0, 0, 0, 1};  // Vec4 defaults w to 1 (include/blitz/mesh.hpp:14)
    float depth = std::numeric_limits<float>::infinity();
    uint16_t material = 0;
    uint8_t covered = 0, visible = 0;
};

struct Pixels {
    uint32_t size = 0;
    std::vector<Pixel> pixels;  // row-major
    bool clipped = false;
};

F3 scale3(F3 a, double s) { return {float(a.x * s), float(a.y * s), float(a.z * s)}; }  // mesh.hpp:76
F3 add3(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
double length3(F3 a) { return std::sqrt(dot(a, a)); }

// src/evaluate.cpp:125-189, attributed instantiation (Raster): conservative
// coverage as above, plus for strictly inside samples perspective-correct
// barycentrics, a depth test against the stored float depth, and the
// winner's interpolated normal (geometric normal when degenerate), colour and
// material.
Pixels rasterize_pixels(const Mesh& m, const Bounds& b, const Camera& c, double screen, uint8_t ss, bool force_two) {
    if (!ss || !std::isfinite(screen) || screen <= 0 || screen > 16384) throw std::invalid_argument("invalid raster extent");
    const uint32_t size = uint32_t(std::ceil(screen + 8)) * ss;
    if (uint64_t(size) * size > kMaxSamples) throw LengthError{};
    Pixels out;
    out.size = size;
    out.pixels.resize(size_t(size) * size);
    struct P { double x, y, z; F3 n; F4 col; };
    std::vector<P> proj(m.vertices);
    const double diameter = b.radius * 2;
    for (uint64_t i = 0; i < m.vertices; ++i) {
        const F3 a = m.vertex(i);
        const double x = double(a.x) - b.center.x, y = double(a.y) - b.center.y, z0 = double(a.z) - b.center.z;
        auto along = [&](F3 axis) { return x * axis.x + y * axis.y + z0 * axis.z; };
        const double z = c.distance - along(c.forward);
        const double scale = c.perspective ? c.focal / z : c.scale;
        P& p = proj[i];
        p = {along(c.right) * scale * ss + size * .5, along(c.up) * scale * ss + size * .5, z, {0, 0, 0}, {1, 1, 1, 1}};
        if (m.normals) p.n = normalized({m.normals[3 * i], m.normals[3 * i + 1], m.normals[3 * i + 2]});
        if (m.colors)  // include/blitz/mesh.hpp:20
            p.col = {m.colors[4 * i] / 255.f, m.colors[4 * i + 1] / 255.f, m.colors[4 * i + 2] / 255.f,
                     m.colors[4 * i + 3] / 255.f};
    }
    for (uint64_t f = 0; f < m.faces(); ++f) {
        const uint32_t ia = m.indices[3 * f], ib = m.indices[3 * f + 1], ic = m.indices[3 * f + 2];
        P a = proj[ia], d = proj[ib], e = proj[ic];
        if (c.perspective && (a.z <= 0 || d.z <= 0 || e.z <= 0)) { out.clipped = true; continue; }
        double area = edge(a.x, a.y, d.x, d.y, e.x, e.y);
        if (std::abs(area) < 1e-16) continue;
        const bool back = area < 0, two = force_two || m.two_sided(f);
        if (back && !two) continue;
        auto delta = [&](uint32_t i) {
            const F3 p = m.vertex(i), p0 = m.vertex(ia);
            return F3{float((double(p.x) - p0.x) / diameter), float((double(p.y) - p0.y) / diameter),
                      float((double(p.z) - p0.z) / diameter)};
        };
        F3 geometric = normalized(cross(delta(ib), delta(ic)));
        if (back) { std::swap(d, e); area = -area; geometric = scale3(geometric, -1); }
        const double minx = std::min({a.x, d.x, e.x}), maxx = std::max({a.x, d.x, e.x});
        const double miny = std::min({a.y, d.y, e.y}), maxy = std::max({a.y, d.y, e.y});
        if (!std::isfinite(minx) || !std::isfinite(maxx) || !std::isfinite(miny) || !std::isfinite(maxy) ||
            minx < 0 || miny < 0 || maxx > size || maxy > size) {
            out.clipped = true;
            return out;
        }
        const int x0 = std::max(0, int(std::floor(minx))), x1 = std::min(int(size) - 1, int(std::floor(maxx)));
        const int y0 = std::max(0, int(std::floor(miny))), y1 = std::min(int(size) - 1, int(std::floor(maxy)));
        const double r0 = .5 * (std::abs(e.x - d.x) + std::abs(e.y - d.y));
        const double r1 = .5 * (std::abs(a.x - e.x) + std::abs(a.y - e.y));
        const double r2 = .5 * (std::abs(d.x - a.x) + std::abs(d.y - a.y));
        for (int y = y0; y <= y1; ++y)
            for (int x = x0; x <= x1; ++x) {
                double u = edge(d.x, d.y, e.x, e.y, x + .5, y + .5), v = edge(e.x, e.y, a.x, a.y, x + .5, y + .5),
                       w = edge(a.x, a.y, d.x, d.y, x + .5, y + .5);
                if (u + r0 < 0 || v + r1 < 0 || w + r2 < 0) continue;
                Pixel& px = out.pixels[size_t(y) * size + x];
                px.covered = 1;
                if (u < 0 || v < 0 || w < 0) continue;
                u /= area;
                v /= area;
                w /= area;
                double depth;
                if (c.perspective) {
                    const double sum = u / a.z + v / d.z + w / e.z;
                    depth = 1 / sum;
                    u = u / a.z / sum;
                    v = v / d.z / sum;
                    w = w / e.z / sum;
                } else {
                    depth = u * a.z + v * d.z + w * e.z;
                }
                depth /= diameter;
                if (depth >= px.depth) continue;
                px.depth = float(depth);
                px.visible = 1;
                px.material = m.material(f);
                px.normal = normalized(add3(add3(scale3(a.n, u), scale3(d.n, v)), scale3(e.n, w)));
                if (length3(px.normal) < .5)
                    px.normal = geometric;
                else if (back && ORACLE_PLANT != 5)
                    px.normal = scale3(px.normal, -1);
                px.color = m.colors ? F4{float(u * a.col.x + v * d.col.x + w * e.col.x),
                                         float(u * a.col.y + v * d.col.y + w * e.col.y),
                                         float(u * a.col.z + v * d.col.z + w * e.col.z), 1}
                                    : F4{1, 1, 1, 1};
            }
    }
    return out;
}

Mask coverage_of(const Pixels& r) {
    Mask m;
    m.size = r.size;
    m.clipped = r.clipped;
    m.words.assign((r.pixels.size() + 63) / 64, 0);
    for (size_t i = 0; i < r.pixels.size(); ++i)
        if (r.pixels[i].covered) m.words[i / 64] |= uint64_t(1) << (i % 64);
    return m;
}

// fdlibm e_acos.c notice, preserved as its licence requires:
// Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
// Developed at SunSoft, a Sun Microsystems, Inc. business.
// Permission to use, copy, modify, and distribute this software is freely
// granted, provided that this notice is preserved.
// fdlibm e_acos.c, in the
// operation sequence of src/metric_angle.hpp:20-41: rational approximation of
// asin about 0 and the half-angle forms near +-1, with the split square root
// for x > 0.5.
double metric_acos(double x) {
    const double pi = 3.14159265358979311600e+00, half = 1.57079632679489655800e+00, lo = 6.12323399573676603587e-17;
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
    const double s = std::sqrt(z);
    if (x < 0) return pi - 2 * (s + (r * s - lo));
    uint64_t bits;
    std::memcpy(&bits, &s, 8);
    bits &= 0xffffffff00000000ull;
    double high;
    std::memcpy(&high, &bits, 8);
    const double correction = (z - high * high) / (s + high);
    return 2 * (high + (r * s + correction));
}

// src/evaluate.cpp:83-95
double sample_cost(const Pixel& a, const Pixel& b, double spatial, const Settings& s) {
    double cost = spatial;
    if (s.profile != 0 && s.normal_weight > 0) {
        const double cosine = std::clamp(dot(a.normal, b.normal) / std::max(1e-30, length3(a.normal) * length3(b.normal)),
                                         -1.0, 1.0);
        const double angle = metric_acos(cosine) * s.normal_weight;
        cost += angle * angle;
    }
    if (s.profile == 2) {
        const double x = double(a.color.x) - b.color.x, y = double(a.color.y) - b.color.y,
                     z = double(a.color.z) - b.color.z;
        cost += s.color_weight * s.color_weight * (x * x + y * y + z * z);
        if (a.material != b.material && ORACLE_PLANT != 6) cost += s.material_weight * s.material_weight;
    }
    return cost;
}

// src/evaluate.cpp:217-270 without the witness: the product-metric directed
// Hausdorff distance between visible samples within a window of the limit.
double attributed_distance(const Pixels& a, const Pixels& b, const Settings& s, uint8_t ss, double limit) {
    if (s.profile == 0 || (s.normal_weight == 0 && (s.profile != 2 || (s.color_weight == 0 && s.material_weight == 0))))
        return 0;
    const double invs2 = 1.0 / (ss * ss), limit2 = limit * limit;
    double maximum = 0;
    const int size = int(a.size);
    auto directed = [&](const Pixels& from, const Pixels& to) {
        const int radius = int(std::min(double(size), std::ceil(limit * ss)));
        for (int y = 0; y < size; ++y)
            for (int x = 0; x < size; ++x) {
                const Pixel& p = from.pixels[size_t(y) * size + x];
                if (!p.visible) continue;
                double best = kInf;
                const Pixel& same = to.pixels[size_t(y) * size + x];
                if (same.visible) best = sample_cost(p, same, 0, s);
                if (best <= 1e-18) continue;
                const int r = int(std::min(double(radius), std::ceil(std::sqrt(std::min(best, limit2)) * ss)));
                for (int yy = std::max(0, y - r); yy <= std::min(size - 1, y + r); ++yy)
                    for (int xx = std::max(0, x - r); xx <= std::min(size - 1, x + r); ++xx) {
                        const Pixel& q = to.pixels[size_t(yy) * size + xx];
                        if (!q.visible) continue;
                        const double dx = x - xx, dy = y - yy, spatial = (dx * dx + dy * dy) * invs2;
                        if (spatial >= best || spatial > limit2) continue;
                        best = std::min(best, sample_cost(p, q, spatial, s));
                    }
                if (best > limit2) return kInf;
                maximum = std::max(maximum, best);
            }
        return maximum;
    };
    if (!std::isfinite(directed(a, b))) return kInf;
    if (!std::isfinite(directed(b, a))) return kInf;
    return std::sqrt(maximum);
}

// src/evaluate.cpp:271-298, attributed profiles.
void finish_view_pixels(Measurement& cur, const Pixels& a, const Pixels& c, const Settings& s, uint8_t ss,
                        double coverage) {
    cur.supersample = ss;
    cur.coverage = coverage;
    cur.coverage_upper = cur.coverage + 2 * std::sqrt(2.0) / ss + 1e-6;
    if (c.clipped || a.clipped) cur.coverage_upper = kInf;
    uint64_t changed = 0, total = 0;
    for (size_t i = 0; i < a.pixels.size(); ++i) {
        changed += a.pixels[i].covered != c.pixels[i].covered;
        total += a.pixels[i].covered || c.pixels[i].covered;
        if (a.pixels[i].visible && c.pixels[i].visible)
            cur.normal_degrees = std::max(
                cur.normal_degrees, metric_acos(std::clamp(dot(a.pixels[i].normal, c.pixels[i].normal), -1.0, 1.0)) *
                                        180 / kPi);
    }
    cur.error = cur.coverage_upper > s.limit ? cur.coverage_upper
                                             : std::max(cur.coverage_upper, attributed_distance(a, c, s, ss, s.limit));
    cur.changed_area = total ? double(changed) / total : 0;
    cur.passed = cur.error <= s.limit && cur.changed_area <= s.max_changed_area;
}

// src/evaluation_settings.hpp:10-27
void validate_settings(const Bounds& b, const Settings& s) {
    if (!(s.screen_size > 0) || !std::isfinite(s.screen_size) || s.screen_size > 16384 || !(b.radius > 0) ||
        !std::isfinite(b.radius) || !std::isfinite(b.center.x) || !std::isfinite(b.center.y) ||
        !std::isfinite(b.center.z) || unsigned(s.profile) > 2 || !(s.limit >= 0) || !std::isfinite(s.limit) ||
        !s.supersample || s.max_supersample < s.supersample || s.max_supersample > 32 ||
        !std::isfinite(s.normal_weight) || !std::isfinite(s.color_weight) || !std::isfinite(s.material_weight) ||
        s.normal_weight < 0 || s.color_weight < 0 || s.material_weight < 0 || s.normal_weight > 1e12 ||
        s.color_weight > 1e12 || s.material_weight > 1e12 || !std::isfinite(s.max_changed_area) ||
        s.max_changed_area < 0 || s.max_changed_area > 1)
        throw std::invalid_argument("invalid evaluation settings");
    if (!s.orthographic && !s.perspective) throw std::invalid_argument("at least one camera required");
}

// src/evaluate.cpp:71-82 (positions, normals, colours, indices, materials,
// sidedness).
bool identical(const Mesh& a, const Mesh& b) {
    auto same = [](const void* x, uint64_t nx, const void* y, uint64_t ny, size_t bytes) {
        return nx == ny && (!nx || x == y || std::memcmp(x, y, nx * bytes) == 0);
    };
    return same(a.positions, a.vertices, b.positions, b.vertices, 12) &&
           same(a.normals, a.normals ? a.vertices : 0, b.normals, b.normals ? b.vertices : 0, 12) &&
           same(a.colors, a.colors ? a.vertices : 0, b.colors, b.colors ? b.vertices : 0, 4) &&
           same(a.indices, a.index_count, b.indices, b.index_count, 4) &&
           same(a.materials, a.material_count, b.materials, b.material_count, 2) &&
           same(a.double_sided, a.double_sided_count, b.double_sided, b.double_sided_count, 1);
}

// src/evaluate.cpp:312-352, coverage profile without the optional cache (the
// cache is result-transparent: tests/coverage_cache.cpp:36-60).
Measurement evaluate(const Mesh& ref, const Mesh& cand, const Bounds& b, const Settings& s,
                     int (*cancelled)(void*), void* user) {
    Measurement result{};
    result.complete = result.passed = 1;
    result.supersample = s.supersample;
    validate_settings(b, s);
    const auto views = make_cameras(b, s.screen_size, s.orthographic, s.perspective, s.rotation_seed);
    if (identical(ref, cand)) return result;
    for (uint32_t v = 0; v < views.size(); ++v) {
        if (cancelled && cancelled(user)) {
            result.complete = result.passed = 0;
            result.cancelled = 1;
            return result;
        }
        Measurement cur{};
        cur.complete = cur.passed = 1;
        cur.worst_view = v;
        for (unsigned ss = s.supersample;; ss = std::min<unsigned>(s.max_supersample, ss * 2)) {
            try {
                if (s.profile == 0 || s.conservative_screen) {
                    const Mask a = rasterize(ref, b, views[v], s.screen_size, uint8_t(ss), s.force_two_sided);
                    const Mask c = rasterize(cand, b, views[v], s.screen_size, uint8_t(ss), s.force_two_sided);
                    finish_view(cur, a, c, s, uint8_t(ss), coverage_distance(a, c, uint8_t(ss)));
                } else {
                    const Pixels a = rasterize_pixels(ref, b, views[v], s.screen_size, uint8_t(ss), s.force_two_sided);
                    const Pixels c = rasterize_pixels(cand, b, views[v], s.screen_size, uint8_t(ss), s.force_two_sided);
                    finish_view_pixels(cur, a, c, s, uint8_t(ss), coverage_distance(coverage_of(a), coverage_of(c),
                                                                                    uint8_t(ss)));
                }
            } catch (const LengthError&) {
                result.complete = result.passed = 0;
                result.resource_limited = 1;
                result.error = kInf;
                return result;
            }
            if (s.conservative_screen || cur.passed || ss >= s.max_supersample ||
                cur.coverage - 2 * std::sqrt(2.0) / ss > s.limit || !std::isfinite(cur.error) ||
                cur.coverage_upper <= s.limit)
                break;
        }
        ++result.views_evaluated;
        if (ORACLE_PLANT == 4 ? cur.error >= result.error : cur.error > result.error) {
            result.worst_view = v;
            result.error = cur.error;
            result.supersample = cur.supersample;
        }
        result.coverage = std::max(result.coverage, cur.coverage);
        result.coverage_upper = std::max(result.coverage_upper, cur.coverage_upper);
        if (cur.changed_area > result.changed_area) {
            result.changed_area = cur.changed_area;
            result.changed_area_worst_view = v;
        }
        result.normal_degrees = std::max(result.normal_degrees, cur.normal_degrees);
        if (!cur.passed) {
            result.passed = result.complete = 0;
            return result;
        }
    }
    return result;
}

// overlap_cuda (src/neural/evaluate.cu:2665-2685): per view, drawn at one
// sample per pixel on the even raster side of src/neural/evaluate.cu:599, the
// drawable faces (setup of src/neural/evaluate.cu:119-176: no stop at an
// out-of-frame face, which only marks the view clipped) are counted per pixel
// when its centre is strictly inside or on a top-left edge
// (src/neural/evaluate.cu:198-200,301-309). A clipped view makes the result
// infinite; otherwise the mean over views of samples / pixels (0 for a view
// with no counted pixel).
double overlap(const Mesh& m, const Bounds& b, double screen, uint16_t ortho, uint16_t persp, uint32_t seed) {
    if (!std::isfinite(screen) || screen <= 0 || screen > 16384) throw std::invalid_argument("invalid raster extent");
    const auto cams = make_cameras(b, screen, ortho, persp, seed);
    if (cams.empty()) throw std::invalid_argument("overlap requires views");
    const uint32_t size = (uint32_t(std::ceil(screen + 8)) + 1) & ~1u;
    if (uint64_t(size) * size > kMaxSamples) throw LengthError{};
    struct P { double x, y, z; };
    auto top_left = [](P a, P c) { return c.y < a.y || (c.y == a.y && c.x < a.x); };
    double total = 0;
    for (const Camera& c : cams) {
        std::vector<P> proj(m.vertices);
        for (uint64_t i = 0; i < m.vertices; ++i) {
            const F3 a = m.vertex(i);
            const double x = double(a.x) - b.center.x, y = double(a.y) - b.center.y, z0 = double(a.z) - b.center.z;
            auto along = [&](F3 axis) { return x * axis.x + y * axis.y + z0 * axis.z; };
            const double z = c.distance - along(c.forward);
            const double scale = c.perspective ? c.focal / z : c.scale;
            proj[i] = {along(c.right) * scale + size * .5, along(c.up) * scale + size * .5, z};
        }
        std::vector<uint32_t> count(size_t(size) * size, 0);
        bool clipped = false;
        for (uint64_t f = 0; f < m.faces(); ++f) {
            P a = proj[m.indices[3 * f]], d = proj[m.indices[3 * f + 1]], e = proj[m.indices[3 * f + 2]];
            if (c.perspective && (a.z <= 0 || d.z <= 0 || e.z <= 0)) { clipped = true; continue; }
            const double area = edge(a.x, a.y, d.x, d.y, e.x, e.y);
            if (std::abs(area) < 1e-16) continue;
            const bool back = area < 0;
            if (back && !m.two_sided(f)) continue;
            if (back) std::swap(d, e);
            const double minx = std::min({a.x, d.x, e.x}), maxx = std::max({a.x, d.x, e.x});
            const double miny = std::min({a.y, d.y, e.y}), maxy = std::max({a.y, d.y, e.y});
            if (!std::isfinite(minx) || !std::isfinite(maxx) || !std::isfinite(miny) || !std::isfinite(maxy) ||
                minx < 0 || miny < 0 || maxx > size || maxy > size) { clipped = true; continue; }
            const int x0 = std::max(0, int(std::floor(minx))), x1 = std::min(int(size) - 1, int(std::floor(maxx)));
            const int y0 = std::max(0, int(std::floor(miny))), y1 = std::min(int(size) - 1, int(std::floor(maxy)));
            const bool t0 = top_left(d, e), t1 = top_left(e, a), t2 = top_left(a, d);
            for (int y = y0; y <= y1; ++y)
                for (int x = x0; x <= x1; ++x) {
                    const double u = edge(d.x, d.y, e.x, e.y, x + .5, y + .5);
                    const double v = edge(e.x, e.y, a.x, a.y, x + .5, y + .5);
                    const double w = edge(a.x, a.y, d.x, d.y, x + .5, y + .5);
                    if (u < 0 || v < 0 || w < 0) continue;
                    if (ORACLE_PLANT == 7 || ((u > 0 || t0) && (v > 0 || t1) && (w > 0 || t2)))
                        ++count[size_t(y) * size + x];
                }
        }
        if (clipped) return kInf;
        uint64_t samples = 0, pixels = 0;
        for (uint32_t k : count)
            if (k) {
                samples += k;
                ++pixels;
            }
        total += pixels ? double(samples) / pixels : 0;
    }
    return total / cams.size();
}

// src/mesh.cpp:77-86
Bounds bounds_of(const Mesh& m) {
    if (!m.vertices) return {{0, 0, 0}, 0};
    F3 lo = m.vertex(0), hi = lo;
    for (uint64_t i = 1; i < m.vertices; ++i) {
        const F3 p = m.vertex(i);
        lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
        hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
    }
    Bounds b{{float((double(lo.x) + hi.x) * .5), float((double(lo.y) + hi.y) * .5), float((double(lo.z) + hi.z) * .5)}, 0};
    for (uint64_t i = 0; i < m.vertices; ++i) {
        const F3 p = m.vertex(i);
        const double x = double(p.x) - b.center.x, y = double(p.y) - b.center.y, z = double(p.z) - b.center.z;
        b.radius = std::max(b.radius, std::sqrt(x * x + y * y + z * z));
    }
    return b;
}

Mesh mesh_of(const float* pos, uint64_t nv, const uint32_t* idx, uint64_t ni, const uint16_t* mat, uint64_t nm,
             const uint8_t* ds, uint64_t nds) {
    return {pos, nv, idx, ni, mat, nm, ds, nds};
}
}  // namespace

extern "C" {
// Status codes: 0 ok, 1 invalid argument, 2 raster over the sample cap.
int oracle_bounds(const float* pos, uint64_t nv, float* center, double* radius) {
    const Bounds b = bounds_of(mesh_of(pos, nv, nullptr, 0, nullptr, 0, nullptr, 0));
    center[0] = b.center.x;
    center[1] = b.center.y;
    center[2] = b.center.z;
    *radius = b.radius;
    return 0;
}

int oracle_cameras(const float* center, double radius, double screen, uint16_t ortho, uint16_t persp, uint32_t seed,
                   Camera* out, uint64_t capacity) {
    const Bounds b{{center[0], center[1], center[2]}, radius};
    const auto cams = make_cameras(b, screen, ortho, persp, seed);
    if (cams.size() > capacity) return 1;
    std::copy(cams.begin(), cams.end(), out);
    return 0;
}

// Row-major packed coverage mask of one view (detail::rasterize_coverage).
int oracle_raster(const float* pos, uint64_t nv, const uint32_t* idx, uint64_t ni, const uint16_t* mat, uint64_t nm,
                  const uint8_t* ds, uint64_t nds, const float* center, double radius, const Camera* cam,
                  double screen, uint8_t ss, uint8_t force_two, uint64_t* words, uint64_t capacity,
                  uint8_t* clipped, uint32_t* size_out) {
    try {
        const Mask r = rasterize(mesh_of(pos, nv, idx, ni, mat, nm, ds, nds), {{center[0], center[1], center[2]}, radius},
                                 *cam, screen, ss, force_two);
        *size_out = r.size;
        *clipped = r.clipped;
        if (r.words.size() > capacity) return 1;
        std::copy(r.words.begin(), r.words.end(), words);
        return 0;
    } catch (const LengthError&) {
        return 2;
    } catch (...) {
        return 1;
    }
}

int oracle_overlap(const float* pos, uint64_t nv, const uint32_t* idx, uint64_t ni, const uint16_t* mat, uint64_t nm,
                   const uint8_t* ds, uint64_t nds, const float* center, double radius, double screen, uint16_t ortho,
                   uint16_t persp, uint32_t seed, double* out) {
    try {
        *out = overlap(mesh_of(pos, nv, idx, ni, mat, nm, ds, nds), {{center[0], center[1], center[2]}, radius}, screen,
                       ortho, persp, seed);
        return 0;
    } catch (const LengthError&) {
        return 2;
    } catch (...) {
        return 1;
    }
}

// Squared EDT of a row-major mask (sites = covered samples), float output.
int oracle_distance_field(const uint64_t* words, uint32_t size, float* out) {
    Mask m;
    m.size = size;
    m.words.assign(words, words + (size_t(size) * size + 63) / 64);
    const auto f = distance_field(m);
    std::copy(f.begin(), f.end(), out);
    return 0;
}

int oracle_coverage_distance(const uint64_t* a, const uint64_t* b, uint32_t size, uint8_t ss, double* out) {
    if (!ss) return 1;
    Mask x, y;
    x.size = y.size = size;
    const size_t n = (size_t(size) * size + 63) / 64;
    x.words.assign(a, a + n);
    y.words.assign(b, b + n);
    *out = coverage_distance(x, y, ss);
    return 0;
}

// Attribute raster of one view (rasterize, src/evaluate.cpp:125-189) as
// row-major arrays: covered, visible, depth, normal xyz, colour rgba, material.
int oracle_raster_pixels(const float* pos, uint64_t nv, const uint32_t* idx, uint64_t ni, const uint16_t* mat,
                         uint64_t nm, const uint8_t* ds, uint64_t nds, const float* normals, const uint8_t* colors,
                         const float* center, double radius, const Camera* cam, double screen, uint8_t ss,
                         uint8_t force_two, uint8_t* covered, uint8_t* visible, float* depth, float* normal,
                         float* color, uint16_t* material, uint64_t capacity, uint8_t* clipped, uint32_t* size_out) {
    try {
        Mesh m = mesh_of(pos, nv, idx, ni, mat, nm, ds, nds);
        m.normals = normals;
        m.colors = colors;
        const Pixels r = rasterize_pixels(m, {{center[0], center[1], center[2]}, radius}, *cam, screen, ss, force_two);
        *size_out = r.size;
        *clipped = r.clipped;
        if (r.pixels.size() > capacity) return 1;
        for (size_t i = 0; i < r.pixels.size(); ++i) {
            const Pixel& p = r.pixels[i];
            covered[i] = p.covered;
            visible[i] = p.visible;
            depth[i] = p.depth;
            normal[3 * i] = p.normal.x;
            normal[3 * i + 1] = p.normal.y;
            normal[3 * i + 2] = p.normal.z;
            color[4 * i] = p.color.x;
            color[4 * i + 1] = p.color.y;
            color[4 * i + 2] = p.color.z;
            color[4 * i + 3] = p.color.w;
            material[i] = p.material;
        }
        return 0;
    } catch (const LengthError&) {
        return 2;
    } catch (...) {
        return 1;
    }
}

double oracle_metric_acos(double x) { return metric_acos(x); }

// evaluate() with normal and colour streams (null when absent).
int oracle_evaluate_streams(const float* rpos, uint64_t rnv, const uint32_t* ridx, uint64_t rni,
                            const uint16_t* rmat, uint64_t rnm, const uint8_t* rds, uint64_t rnds,
                            const float* rnormals, const uint8_t* rcolors, const float* cpos, uint64_t cnv,
                            const uint32_t* cidx, uint64_t cni, const uint16_t* cmat, uint64_t cnm,
                            const uint8_t* cds, uint64_t cnds, const float* cnormals, const uint8_t* ccolors,
                            const float* center, double radius, const Settings* s, Measurement* out,
                            int (*cancelled)(void*), void* user) {
    try {
        Mesh r = mesh_of(rpos, rnv, ridx, rni, rmat, rnm, rds, rnds), c = mesh_of(cpos, cnv, cidx, cni, cmat, cnm, cds, cnds);
        r.normals = rnormals;
        r.colors = rcolors;
        c.normals = cnormals;
        c.colors = ccolors;
        *out = evaluate(r, c, {{center[0], center[1], center[2]}, radius}, *s, cancelled, user);
        return 0;
    } catch (...) {
        return 1;
    }
}

int oracle_evaluate(const float* rpos, uint64_t rnv, const uint32_t* ridx, uint64_t rni, const uint16_t* rmat,
                    uint64_t rnm, const uint8_t* rds, uint64_t rnds, const float* cpos, uint64_t cnv,
                    const uint32_t* cidx, uint64_t cni, const uint16_t* cmat, uint64_t cnm, const uint8_t* cds,
                    uint64_t cnds, const float* center, double radius, const Settings* s, Measurement* out,
                    int (*cancelled)(void*), void* user) {
    try {
        *out = evaluate(mesh_of(rpos, rnv, ridx, rni, rmat, rnm, rds, rnds),
                        mesh_of(cpos, cnv, cidx, cni, cmat, cnm, cds, cnds), {{center[0], center[1], center[2]}, radius},
                        *s, cancelled, user);
        return 0;
    } catch (...) {
        return 1;
    }
}
}
