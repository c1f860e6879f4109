// Classical quadric edge-collapse reducer with reused vertices.
//
// The counterpart of the reference's classical proposer (blitz::reduce with
// OutputMode::Reuse, src/simplify.cpp:43-367): FP64 plane quadrics
// (Garland & Heckbert 1997) accumulated per vertex, area weighted, in
// coordinates normalised by the bounding sphere; each collapse moves one
// endpoint onto the other, so the output indexes the input vertex array.
// Unlike the reference's pass-based sweep, collapses are taken greedily from a
// cost heap (the textbook order), with the link condition and a normal-flip
// guard keeping the surface manifold. Used to build seeded audit workloads and
// the plain classical order; its output is audited like any proposal.
#include "simplify.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <queue>
#include <stdexcept>
#include <unordered_map>
#include <vector>

namespace blitz_hip {
namespace {
struct Quadric {
    double a[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    void plane(double x, double y, double z, double w, double weight) {
        const double v[4] = {x, y, z, w};
        int k = 0;
        for (int i = 0; i < 4; ++i)
            for (int j = i; j < 4; ++j) a[k++] += v[i] * v[j] * weight;
    }
    void add(const Quadric& o) {
        for (int i = 0; i < 10; ++i) a[i] += o.a[i];
    }
    double cost(const double* p) const {
        const double x = p[0], y = p[1], z = p[2];
        return std::max(0.0, a[0] * x * x + 2 * a[1] * x * y + 2 * a[2] * x * z + 2 * a[3] * x + a[4] * y * y +
                                 2 * a[5] * y * z + 2 * a[6] * y + a[7] * z * z + 2 * a[8] * z + a[9]);
    }
};

struct Entry {
    double cost;
    uint32_t from, to, stamp_from, stamp_to;
    bool operator>(const Entry& o) const {
        if (cost != o.cost) return cost > o.cost;
        if (from != o.from) return from > o.from;
        return to > o.to;
    }
};

void cross(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
}  // namespace

std::vector<uint32_t> reduce_reuse(const float* xyz, size_t nv, const uint32_t* indices, size_t nf, size_t target,
                                   uint64_t* collapses) {
    if (collapses) *collapses = 0;
    std::vector<uint32_t> face(indices, indices + 3 * nf);
    if (nf <= target || nv == 0) return face;
    // Normalise like the reference (p - center) / diameter for conditioning.
    double lo[3] = {xyz[0], xyz[1], xyz[2]}, hi[3] = {xyz[0], xyz[1], xyz[2]};
    for (size_t i = 1; i < nv; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], double(xyz[3 * i + k]));
            hi[k] = std::max(hi[k], double(xyz[3 * i + k]));
        }
    double center[3], radius = 0;
    for (int k = 0; k < 3; ++k) center[k] = .5 * (lo[k] + hi[k]);
    for (size_t i = 0; i < nv; ++i) {
        double d2 = 0;
        for (int k = 0; k < 3; ++k) d2 += (xyz[3 * i + k] - center[k]) * (xyz[3 * i + k] - center[k]);
        radius = std::max(radius, std::sqrt(d2));
    }
    const double scale = radius > 0 ? 2 * radius : 1;
    std::vector<double> p(3 * nv);
    for (size_t i = 0; i < nv; ++i)
        for (int k = 0; k < 3; ++k) p[3 * i + k] = (double(xyz[3 * i + k]) - center[k]) / scale;

    std::vector<Quadric> q(nv);
    std::vector<std::vector<uint32_t>> faces_of(nv);
    std::vector<uint8_t> alive_face(nf, 1);
    size_t live = 0;
    for (size_t f = 0; f < nf; ++f) {
        const uint32_t* t = &face[3 * f];
        if (t[0] == t[1] || t[1] == t[2] || t[0] == t[2]) {
            alive_face[f] = 0;
            continue;
        }
        double e1[3], e2[3], n[3];
        for (int k = 0; k < 3; ++k) {
            e1[k] = p[3 * t[1] + k] - p[3 * t[0] + k];
            e2[k] = p[3 * t[2] + k] - p[3 * t[0] + k];
        }
        cross(e1, e2, n);
        const double area = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (area > 1e-24) {
            for (int k = 0; k < 3; ++k) n[k] /= area;
            const double w = -(n[0] * p[3 * t[0]] + n[1] * p[3 * t[0] + 1] + n[2] * p[3 * t[0] + 2]);
            for (int j = 0; j < 3; ++j) q[t[j]].plane(n[0], n[1], n[2], w, area);
        }
        for (int j = 0; j < 3; ++j) faces_of[t[j]].push_back(uint32_t(f));
        ++live;
    }
    // Boundary edges (one incident face) get a perpendicular constraint plane
    // so open borders keep their silhouette.
    {
        std::unordered_map<uint64_t, uint32_t> count;
        count.reserve(nf * 3);
        for (size_t f = 0; f < nf; ++f) {
            if (!alive_face[f]) continue;
            for (int j = 0; j < 3; ++j) {
                uint32_t a = face[3 * f + j], b = face[3 * f + (j + 1) % 3];
                if (a > b) std::swap(a, b);
                ++count[(uint64_t(a) << 32) | b];
            }
        }
        for (size_t f = 0; f < nf; ++f) {
            if (!alive_face[f]) continue;
            const uint32_t* t = &face[3 * f];
            double e1[3], e2[3], n[3];
            for (int k = 0; k < 3; ++k) {
                e1[k] = p[3 * t[1] + k] - p[3 * t[0] + k];
                e2[k] = p[3 * t[2] + k] - p[3 * t[0] + k];
            }
            cross(e1, e2, n);
            for (int j = 0; j < 3; ++j) {
                uint32_t a = t[j], b = t[(j + 1) % 3];
                const uint64_t key = a < b ? (uint64_t(a) << 32) | b : (uint64_t(b) << 32) | a;
                if (count[key] != 1) continue;
                double e[3] = {p[3 * b] - p[3 * a], p[3 * b + 1] - p[3 * a + 1], p[3 * b + 2] - p[3 * a + 2]}, side[3];
                cross(n, e, side);
                const double len = std::sqrt(side[0] * side[0] + side[1] * side[1] + side[2] * side[2]);
                if (len <= 1e-24) continue;
                for (int k = 0; k < 3; ++k) side[k] /= len;
                const double w = -(side[0] * p[3 * a] + side[1] * p[3 * a + 1] + side[2] * p[3 * a + 2]);
                const double weight = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
                q[a].plane(side[0], side[1], side[2], w, weight);
                q[b].plane(side[0], side[1], side[2], w, weight);
            }
        }
    }
    std::vector<uint32_t> stamp(nv, 0);
    std::vector<uint8_t> alive(nv, 1);
    std::priority_queue<Entry, std::vector<Entry>, std::greater<Entry>> heap;
    auto push_edge = [&](uint32_t u, uint32_t v) {
        Quadric s = q[u];
        s.add(q[v]);
        const double cu = s.cost(&p[3 * v]);  // u moves onto v
        const double cv = s.cost(&p[3 * u]);  // v moves onto u
        if (cu <= cv)
            heap.push({cu, u, v, stamp[u], stamp[v]});
        else
            heap.push({cv, v, u, stamp[v], stamp[u]});
    };
    {
        std::vector<uint64_t> keys;
        keys.reserve(3 * nf);
        for (size_t f = 0; f < nf; ++f) {
            if (!alive_face[f]) continue;
            for (int j = 0; j < 3; ++j) {
                const uint32_t a = face[3 * f + j], b = face[3 * f + (j + 1) % 3];
                keys.push_back(a < b ? (uint64_t(a) << 32) | b : (uint64_t(b) << 32) | a);
            }
        }
        std::sort(keys.begin(), keys.end());
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
        for (uint64_t k : keys) push_edge(uint32_t(k >> 32), uint32_t(k));
    }
    std::vector<uint32_t> mark(nv, 0);
    uint32_t epoch = 0;
    while (live > target && !heap.empty()) {
        const Entry e = heap.top();
        heap.pop();
        const uint32_t u = e.from, v = e.to;
        if (!alive[u] || !alive[v] || stamp[u] != e.stamp_from || stamp[v] != e.stamp_to) continue;
        // Link condition: the common neighbours of u and v are exactly the
        // apexes of the faces sharing edge uv.
        ++epoch;
        for (uint32_t f : faces_of[u])
            if (alive_face[f])
                for (int k = 0; k < 3; ++k) mark[face[3 * f + k]] = epoch;
        size_t common = 0, shared = 0;
        ++epoch;
        for (uint32_t f : faces_of[v]) {
            if (!alive_face[f]) continue;
            bool has_u = false;
            for (int k = 0; k < 3; ++k) has_u |= face[3 * f + k] == u;
            shared += has_u;
            for (int k = 0; k < 3; ++k) {
                const uint32_t w = face[3 * f + k];
                if (w == u || w == v) continue;
                if (mark[w] == epoch - 1) {
                    mark[w] = epoch;
                    ++common;
                }
            }
        }
        if (shared == 0 || common != shared) continue;
        // Normal-flip guard on faces that move with u.
        bool ok = true;
        for (uint32_t f : faces_of[u]) {
            if (!alive_face[f]) continue;
            const uint32_t* t = &face[3 * f];
            if (t[0] == v || t[1] == v || t[2] == v) continue;
            double before[3], after[3], e1[3], e2[3];
            const double* pa[3] = {&p[3 * t[0]], &p[3 * t[1]], &p[3 * t[2]]};
            for (int k = 0; k < 3; ++k) {
                e1[k] = pa[1][k] - pa[0][k];
                e2[k] = pa[2][k] - pa[0][k];
            }
            cross(e1, e2, before);
            const double* pb[3];
            for (int j = 0; j < 3; ++j) pb[j] = t[j] == u ? &p[3 * v] : pa[j];
            for (int k = 0; k < 3; ++k) {
                e1[k] = pb[1][k] - pb[0][k];
                e2[k] = pb[2][k] - pb[0][k];
            }
            cross(e1, e2, after);
            const double nb = std::sqrt(before[0] * before[0] + before[1] * before[1] + before[2] * before[2]);
            const double na = std::sqrt(after[0] * after[0] + after[1] * after[1] + after[2] * after[2]);
            const double d = before[0] * after[0] + before[1] * after[1] + before[2] * after[2];
            if (!(na > 1e-30) || d < 0.2 * na * nb) {
                ok = false;
                break;
            }
        }
        if (!ok) continue;
        // Collapse u onto v.
        for (uint32_t f : faces_of[u]) {
            if (!alive_face[f]) continue;
            uint32_t* t = &face[3 * f];
            if (t[0] == v || t[1] == v || t[2] == v) {
                alive_face[f] = 0;
                --live;
                continue;
            }
            for (int k = 0; k < 3; ++k)
                if (t[k] == u) t[k] = v;
            faces_of[v].push_back(f);
        }
        faces_of[u].clear();
        alive[u] = 0;
        q[v].add(q[u]);
        ++stamp[v];
        if (collapses) ++*collapses;
        // Compact v's face list and requeue its edges (entries holding v's old
        // stamp are now stale; edges not touching v keep their cost).
        auto& fv = faces_of[v];
        fv.erase(std::remove_if(fv.begin(), fv.end(), [&](uint32_t f) { return !alive_face[f]; }), fv.end());
        ++epoch;
        for (uint32_t f : fv)
            for (int k = 0; k < 3; ++k) {
                const uint32_t w = face[3 * f + k];
                if (w != v && mark[w] != epoch) {
                    mark[w] = epoch;
                    push_edge(std::min(v, w), std::max(v, w));
                }
            }
    }
    std::vector<uint32_t> out;
    out.reserve(3 * live);
    for (size_t f = 0; f < nf; ++f)
        if (alive_face[f]) out.insert(out.end(), &face[3 * f], &face[3 * f] + 3);
    return out;
}

}  // namespace blitz_hip
