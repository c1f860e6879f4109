// Float 3-vectors with the reference's Vec3 arithmetic on the device: float
// differences and cross products (each product and difference rounded to
// float, no contraction), dot products and lengths in double
// (include/blitz/mesh.hpp Vec3 operators; src/neural/action.cpp:355-370 and
// src/neural/action_gpu.cu:809-844 rank with these).
#pragma once
#include <hip/hip_runtime.h>

#include "audit_types.hpp"

namespace blitz_hip {

__device__ __forceinline__ F3 sub3(float4 a, float4 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 cross3(F3 a, F3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double dot3(F3 a, F3 b) {
    return double(a.x) * b.x + double(a.y) * b.y + double(a.z) * b.z;
}
__device__ __forceinline__ double len3(F3 a) { return sqrt(dot3(a, a)); }
__device__ __forceinline__ F3 sub3(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 add3(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
// Vec3 * double: each product in double, rounded to float.
__device__ __forceinline__ F3 scale3(F3 a, double s) { return {float(a.x * s), float(a.y * s), float(a.z * s)}; }
// normalized(): a * (1 / length) when the length is positive, else zero.
__device__ __forceinline__ F3 norm3(F3 a) {
    const double l = len3(a);
    return l > 0 ? scale3(a, 1 / l) : F3{0, 0, 0};
}
__device__ __forceinline__ F3 xyz(float4 p) { return {p.x, p.y, p.z}; }

// Ranking scores of the plain orders (src/neural/action_gpu.cu:822-840,
// CPU ActionState::teacher_cost src/neural/action.cpp:355-370): shortest
// edge -|to - from| / scale, and current plane: minus the sum over the faces
// around `from` of (plane distance of `to` / scale)^2 times the face's
// |cross| / scale^2. `scale` is the input's diameter (at least 1e-20).
__device__ __forceinline__ float shortest_edge_score(const float4* positions, uint32_t from, uint32_t to,
                                                     double scale) {
    const F3 d = sub3(positions[to], positions[from]);
    const double s = 1 / scale;
    const F3 e{float(d.x * s), float(d.y * s), float(d.z * s)};
    return -float(len3(e));
}
__device__ __forceinline__ float current_plane_score(const float4* positions, const uint32_t* indices,
                                                     const uint32_t* face_offsets, const uint32_t* faces,
                                                     uint32_t from, uint32_t to, double scale) {
    double cost = 0;
    for (uint32_t k = face_offsets[from]; k < face_offsets[from + 1]; ++k) {
        const uint32_t f = faces[k];
        const float4 p = positions[indices[3 * f]], q = positions[indices[3 * f + 1]], r = positions[indices[3 * f + 2]];
        const F3 normal = cross3(sub3(q, p), sub3(r, p));
        const double weight = len3(normal);
        const F3 unit = weight > 0 ? F3{float(normal.x * (1 / weight)), float(normal.y * (1 / weight)),
                                         float(normal.z * (1 / weight))}
                                   : F3{0, 0, 0};
        const double distance = dot3(unit, sub3(positions[to], p)) / scale;
        cost += distance * distance * weight / (scale * scale);
    }
    return float(-cost);
}

}  // namespace blitz_hip
