"""Meshes for the action-feature parity tests: each exercises some of the
graph channels of src/neural/graph.cpp (seams from split UVs and normals,
colours, tangents, two materials, open boundaries, closed surfaces)."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
from blitzremesher_amd import meshes  # noqa: E402
from blitzremesher_amd.audit import Mesh  # noqa: E402

CONDITIONS = np.array([0.75, 0.0625, 0.0625, 0.3, 0.25, 0.25, 1.0, 0.5], np.float32)


def _tangents(n: int, seed: int) -> np.ndarray:
    t = np.stack([meshes.normal(seed, 21, n), meshes.normal(seed, 22, n), meshes.normal(seed, 23, n),
                  np.where(meshes.uniform(seed, 24, n) < 0.5, -1.0, 1.0)], -1)
    return t.astype(np.float32)


def seamed(mesh: Mesh, seed: int) -> tuple[Mesh, np.ndarray, np.ndarray]:
    """`mesh` with the faces right of x = median moved onto duplicate vertices
    whose UVs (and normals) differ: every position on the split line holds two
    attribute classes (a seam). Two materials, colours and tangents."""
    p = mesh.positions
    n = len(p)
    idx = mesh.indices.astype(np.int64).copy()
    centre = p[idx].mean(1)[:, 0]
    right = centre > np.median(centre)
    idx[right] += n
    pos = np.concatenate([p, p])
    uv = np.concatenate([p[:, :2] * 2.0, p[:, :2] * 2.0 + 0.5]).astype(np.float32)
    base = meshes.with_attributes(mesh, seed)
    normals = np.concatenate([base.normals, base.normals[:, ::-1]]).astype(np.float32)
    colors = np.concatenate([base.colors, base.colors[::-1]])
    mats = ((np.arange(len(idx)) // 7) % 2).astype(np.uint16)
    out = Mesh(pos, idx.astype(np.uint32), mats, None, normals, colors)
    return out, uv, _tangents(2 * n, seed)


def cases() -> dict:
    """name -> (mesh, uv or None, tangents or None)."""
    grid = meshes.jitter(meshes.grid(6, double_sided=False), 0.01, 3)
    ico = meshes.icosphere(1, 0.02, 5)
    out = {
        "seamed_grid": seamed(grid, 4),
        "icosphere_attributes": (meshes.with_attributes(ico, 6), None, None),
        "grid_uv": (grid, (grid.positions[:, :2] * 3.0).astype(np.float32), None),
        "torus_materials": (Mesh(meshes.small_torus().positions, meshes.small_torus().indices,
                                 (np.arange(160) % 3).astype(np.uint16)), None,
                            _tangents(80, 9)),
    }
    return out
