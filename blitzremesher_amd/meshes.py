"""Seeded synthetic meshes for the audit and scorer workloads.

Randomness is a counter-based splitmix64 stream (value i of stream s under
seed k is splitmix64(k ^ splitmix64(s) ^ i * golden)), so every generator is
vectorised, deterministic across processes and independent of call order.
Gaussian noise uses Box-Muller on two uniform streams.
"""
from __future__ import annotations

import numpy as np

from .audit import Mesh

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def _splitmix(x: np.ndarray) -> np.ndarray:
    x = (x + _GOLDEN).astype(np.uint64)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """n uniform doubles in [0, 1)."""
    with np.errstate(over="ignore"):
        key = _splitmix(np.array([seed & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64))[0] ^ \
            _splitmix(np.array([stream], dtype=np.uint64))[0]
        bits = _splitmix(np.arange(n, dtype=np.uint64) * _GOLDEN ^ key)
    return (bits >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    u1 = uniform(seed, 2 * stream, n)
    u2 = uniform(seed, 2 * stream + 1, n)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def grid(n: int, double_sided: bool = True) -> Mesh:
    """Unit square lattice in the z = 0 plane, two triangles per cell (the
    reference test fixture, tests/contracts.cpp:11-16)."""
    ys, xs = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    pos = np.stack([xs.astype(np.float32) / np.float32(n), ys.astype(np.float32) / np.float32(n),
                    np.zeros_like(xs, dtype=np.float32)], axis=-1).reshape(-1, 3)
    y, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    a = (y * (n + 1) + x).reshape(-1)
    b, c = a + 1, a + n + 1
    d = c + 1
    idx = np.stack([a, b, d, a, d, c], axis=-1).reshape(-1, 3)
    return Mesh(pos, idx, double_sided=np.array([1], np.uint8) if double_sided else None)


def small_torus() -> Mesh:
    """10 x 8 torus of tests/contracts.cpp:22-33."""
    major, minor = 10, 8
    pos = []
    for i in range(major):
        for j in range(minor):
            a = i * 6.283185307179586 / major
            b = j * 6.283185307179586 / minor
            pos.append([np.float32((2 + .6 * np.cos(b)) * np.cos(a)), np.float32((2 + .6 * np.cos(b)) * np.sin(a)),
                        np.float32(.6 * np.sin(b))])
    idx = []
    for i in range(major):
        for j in range(minor):
            a = i * minor + j
            b = ((i + 1) % major) * minor + j
            c = ((i + 1) % major) * minor + (j + 1) % minor
            d = i * minor + (j + 1) % minor
            idx += [[a, b, c], [a, c, d]]
    return Mesh(np.array(pos, np.float32), np.array(idx, np.uint32))


def four_tiles() -> Mesh:
    """Four separated unit squares (tests/contracts.cpp:34-42)."""
    pos, idx = [], []
    for y in (0.0, 2.0):
        for x in (0.0, 2.0):
            base = len(pos)
            pos += [[x, y, 0], [x + 1, y, 0], [x + 1, y + 1, 0], [x, y + 1, 0]]
            idx += [[base, base + 1, base + 2], [base, base + 2, base + 3]]
    return Mesh(np.array(pos, np.float32), np.array(idx, np.uint32), double_sided=np.array([1], np.uint8))


def _icosahedron() -> tuple[np.ndarray, np.ndarray]:
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def icosphere(subdivisions: int = 5, noise: float = 0.02, seed: int = 1) -> Mesh:
    """Unit icosphere subdivided `subdivisions` times (10,242 V / 20,480 F at 5),
    each vertex scaled radially by 1 + N(0, noise)."""
    v, f = _icosahedron()
    for _ in range(subdivisions):
        edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = np.sort(edges, axis=1)
        uniq, inverse = np.unique(key, axis=0, return_inverse=True)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = inverse.reshape(3, -1) + len(v)
        v = np.concatenate([v, mid])
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        ab, bc, ca = m[0], m[1], m[2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                            np.stack([ab, bc, ca], 1)])
    if noise:
        v = v * (1.0 + noise * normal(seed, 0, len(v)))[:, None]
    return Mesh(v.astype(np.float32), f.astype(np.uint32))


def _value_noise(x: np.ndarray, y: np.ndarray, seed: int, octave: int) -> np.ndarray:
    xi, yi = np.floor(x).astype(np.int64), np.floor(y).astype(np.int64)
    fx, fy = x - xi, y - yi
    sx, sy = fx * fx * (3 - 2 * fx), fy * fy * (3 - 2 * fy)

    def lattice(i, j):
        with np.errstate(over="ignore"):
            k = (i.astype(np.uint64) * np.uint64(0x8DA6B343)) ^ (j.astype(np.uint64) * np.uint64(0xD8163841)) ^ \
                np.uint64((seed * 0x9E3779B1 + octave * 0x85EBCA77) & 0xFFFFFFFFFFFFFFFF)
            h = _splitmix(k)
        return (h >> np.uint64(11)).astype(np.float64) * (2.0 / (1 << 53)) - 1.0

    n00, n10, n01, n11 = lattice(xi, yi), lattice(xi + 1, yi), lattice(xi, yi + 1), lattice(xi + 1, yi + 1)
    return (n00 * (1 - sx) + n10 * sx) * (1 - sy) + (n01 * (1 - sx) + n11 * sx) * sy


def fbm(x: np.ndarray, y: np.ndarray, seed: int, octaves: int = 6, lacunarity: float = 2.0, gain: float = 0.5,
        frequency: float = 4.0) -> np.ndarray:
    total = np.zeros_like(x)
    amplitude, norm = 1.0, 0.0
    for o in range(octaves):
        total += amplitude * _value_noise(x * frequency, y * frequency, seed, o)
        norm += amplitude
        amplitude *= gain
        frequency *= lacunarity
    return total / norm


def heightfield(n: int = 512, amplitude: float = 0.1, seed: int = 2) -> Mesh:
    """n x n lattice on [0,1]^2 with z = amplitude * fBm (6 octaves, lacunarity 2,
    gain 0.5): n^2 vertices and 2 (n-1)^2 triangles (262,144 / 522,242 at 512)."""
    ys, xs = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    u, v = xs / (n - 1), ys / (n - 1)
    z = amplitude * fbm(u, v, seed)
    pos = np.stack([u, v, z], -1).reshape(-1, 3).astype(np.float32)
    y, x = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    a = (y * n + x).reshape(-1)
    b, c = a + 1, a + n
    d = c + 1
    idx = np.stack([a, b, d, a, d, c], -1).reshape(-1, 3)
    return Mesh(pos, idx)


def torus(major_radius: float = 1.0, minor_radius: float = 0.3, major: int = 500, minor: int = 200,
          noise: float = 0.01, seed: int = 3) -> Mesh:
    """Closed torus, major x minor quads (100,000 V / 200,000 F at 500 x 200),
    vertices displaced along the normal by N(0, noise)."""
    i, j = np.meshgrid(np.arange(major), np.arange(minor), indexing="ij")
    a = i * 2 * np.pi / major
    b = j * 2 * np.pi / minor
    nrm = np.stack([np.cos(b) * np.cos(a), np.cos(b) * np.sin(a), np.sin(b)], -1)
    center = np.stack([major_radius * np.cos(a), major_radius * np.sin(a), np.zeros_like(a)], -1)
    r = minor_radius + (noise * normal(seed, 0, major * minor).reshape(major, minor) if noise else 0.0)
    pos = (center + nrm * np.asarray(r)[..., None]).reshape(-1, 3).astype(np.float32)
    i, j = i.reshape(-1), j.reshape(-1)
    v00 = i * minor + j
    v10 = ((i + 1) % major) * minor + j
    v11 = ((i + 1) % major) * minor + (j + 1) % minor
    v01 = i * minor + (j + 1) % minor
    idx = np.stack([v00, v10, v11, v00, v11, v01], -1).reshape(-1, 3)
    return Mesh(pos, idx)


def extruded_box(n: int = 16, height: float = 0.5, seed: int = 7) -> Mesh:
    """Closed box from an n x n top grid extruded downwards, with a seeded
    jitter of the top surface."""
    top = grid(n, double_sided=False)
    p = top.positions.astype(np.float64)
    p[:, 2] = height + 0.02 * height * normal(seed, 0, len(p))
    bottom = p.copy()
    bottom[:, 2] = 0.0
    nv = len(p)
    faces = [top.indices, top.indices[:, ::-1] + nv]
    side = []
    ring = [y * (n + 1) for y in range(n + 1)] + [n + y * (n + 1) for y in range(n + 1)]
    boundary = []
    for x in range(n):
        boundary.append((x, x + 1))
        boundary.append(((n) * (n + 1) + x + 1, (n) * (n + 1) + x))
    for y in range(n):
        boundary.append(((y + 1) * (n + 1), y * (n + 1)))
        boundary.append((y * (n + 1) + n, (y + 1) * (n + 1) + n))
    del ring
    for a, b in boundary:
        side += [[b, a, a + nv], [b, a + nv, b + nv]]
    faces.append(np.array(side, np.uint32))
    return Mesh(np.concatenate([p, bottom]).astype(np.float32), np.concatenate(faces))


def jitter(mesh: Mesh, sigma: float, seed: int, fraction: float = 1.0) -> Mesh:
    """Copy with a seeded subset of vertices moved by N(0, sigma) per axis."""
    n = len(mesh.positions)
    pick = uniform(seed, 100, n) < fraction
    d = np.stack([normal(seed, 1, n), normal(seed, 2, n), normal(seed, 3, n)], -1) * sigma
    pos = mesh.positions.astype(np.float64) + d * pick[:, None]
    return Mesh(pos.astype(np.float32), mesh.indices.copy(), mesh.materials, mesh.double_sided)


def drop_faces(mesh: Mesh, fraction: float, seed: int) -> Mesh:
    """Copy without a seeded subset of faces (shares the vertex array)."""
    keep = uniform(seed, 200, mesh.triangles) >= fraction
    mats = mesh.materials[keep] if mesh.materials is not None else None
    return Mesh(mesh.positions, mesh.indices[keep], mats, mesh.double_sided)


def with_attributes(mesh: Mesh, seed: int, normal_noise: float = 0.05, colors: bool = True) -> Mesh:
    """Copy with per-vertex normals (area-weighted face normals plus seeded
    noise, unnormalised) and seeded RGBA8 colours that vary smoothly with the
    position."""
    p = mesh.positions.astype(np.float64)
    f = mesh.indices.astype(np.int64)
    fn = np.cross(p[f[:, 1]] - p[f[:, 0]], p[f[:, 2]] - p[f[:, 0]])
    n = np.zeros_like(p)
    for k in range(3):
        np.add.at(n, f[:, k], fn)
    count = len(p)
    n += np.stack([normal(seed, 11, count), normal(seed, 12, count), normal(seed, 13, count)], -1) * \
        normal_noise * np.linalg.norm(n, axis=1, keepdims=True).clip(1e-12)
    col = None
    if colors:
        phase = uniform(seed, 14, 3) * 6.283185307179586
        rgb = 0.5 + 0.5 * np.sin(p * 3.0 + phase)
        col = np.concatenate([np.round(rgb * 255), np.full((count, 1), 255)], 1).astype(np.uint8)
    return Mesh(mesh.positions, mesh.indices, mesh.materials, mesh.double_sided, n.astype(np.float32), col)


def uv_sphere(rings: int, segments: int, noise: float = 0.02, seed: int = 1) -> Mesh:
    """Closed unit latitude-longitude sphere: `rings` latitude bands and
    `segments` meridians (segments (rings - 1) + 2 vertices, 2 segments
    (rings - 1) triangles, fans at both poles), each vertex scaled radially
    by 1 + N(0, noise)."""
    if rings < 2 or segments < 3:
        raise ValueError("a sphere needs at least 2 rings and 3 segments")
    theta = np.arange(1, rings) * np.pi / rings
    phi = np.arange(segments) * 2 * np.pi / segments
    t, p = np.meshgrid(theta, phi, indexing="ij")
    body = np.stack([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)], -1).reshape(-1, 3)
    v = np.concatenate([body, [[0.0, 0.0, 1.0], [0.0, 0.0, -1.0]]])
    north, south = len(body), len(body) + 1
    r, s = np.meshgrid(np.arange(rings - 2), np.arange(segments), indexing="ij")
    a = r * segments + s
    b = r * segments + (s + 1) % segments
    c, d = a + segments, b + segments
    quads = np.stack([a, c, d, a, d, b], -1).reshape(-1, 3)
    s = np.arange(segments)
    top = np.stack([np.full(segments, north), s, (s + 1) % segments], -1)
    last = (rings - 2) * segments
    bottom = np.stack([np.full(segments, south), last + (s + 1) % segments, last + s], -1)
    if noise:
        v = v * (1.0 + noise * normal(seed, 0, len(v)))[:, None]
    return Mesh(v.astype(np.float32), np.concatenate([top, quads, bottom]).astype(np.uint32))


TRAINING_FAMILIES = ("sphere", "torus", "extruded_box", "heightfield")


def training_mesh(family: str, triangles: int, seed: int) -> Mesh:
    """One mesh of `family` with about `triangles` triangles (within a few
    percent): a noisy latitude-longitude sphere (2 s (r - 1) triangles,
    s = 2 r), a torus with normal noise (major = 2.5 minor), an extruded
    box with a jittered top (4 n^2 + 8 n) or an fBm heightfield
    (2 (n - 1)^2)."""
    t = max(int(triangles), 64)
    if family == "sphere":
        r = max(2, int(round((1 + np.sqrt(1 + 4 * t / 4.0)) / 2)))
        return uv_sphere(r, 2 * r, 0.02, seed)
    if family == "torus":
        minor = max(3, int(round(np.sqrt(t / 5.0))))
        return torus(1.0, 0.3, max(3, int(round(2.5 * minor))), minor, 0.01, seed)
    if family == "extruded_box":
        return extruded_box(max(2, int(round(np.sqrt(1 + t / 4.0) - 1))), 0.5, seed)
    if family == "heightfield":
        return heightfield(max(3, int(round(np.sqrt(t / 2.0))) + 1), 0.1, seed)
    raise ValueError(f"unknown mesh family {family!r}")


def training_set(count: int = 64, low: int = 2000, high: int = 100000, seed: int = 5) -> list[tuple[str, Mesh]]:
    """`count` meshes of uneven size for scorer training: triangle counts
    log-uniform in [low, high] (splitmix stream 40 of `seed`), families in
    turn sphere, torus, extruded box, heightfield, each mesh with its own
    noise seed."""
    if count < 1 or not 0 < low <= high:
        raise ValueError("invalid training set size")
    u = uniform(seed, 40, count)
    sizes = np.exp(np.log(low) + u * (np.log(high) - np.log(low)))
    return [(TRAINING_FAMILIES[i % 4], training_mesh(TRAINING_FAMILIES[i % 4], int(sizes[i]), seed * 1000 + i))
            for i in range(count)]
