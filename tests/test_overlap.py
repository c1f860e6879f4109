"""Overlap under the top-left fill rule (overlap_cuda,
src/neural/evaluate.cu:2665-2685; a tiebreak of the neural generator,
src/neural/generate.cpp:563-570).

The reference computes it on CUDA only, so the oracle's restatement is pinned
by what the fill rule fixes: a closed convex mesh drawn one-sided covers every
counted pixel exactly once (1.0 in every view, orthographic or perspective),
the same mesh drawn two-sided or with its faces duplicated exactly twice, and
a view with a vertex out of frame makes the measure infinite; a grid whose
edges run through pixel centres also measures exactly 1.0, so a planted
mistake (edges counted without the top-left rule) fails. The GPU result
equals the oracle's exactly.
"""
from __future__ import annotations

import numpy as np
import pytest

from blitzremesher_amd import Mesh, meshes
from oracle_lib import Oracle


def convex():
    return meshes.icosphere(3, 0.0, 1)


def pixel_grid():
    """A two-sided 16x16 quad grid in the plane x = 0 whose vertices project,
    for the single orthographic view of seed 0 (forward +x, right -z, up +y)
    with center 0, radius 28 and screen 56 (scale 1, raster side 64), onto
    half-integer raster coordinates: every vertical and diagonal edge passes
    through pixel centres."""
    n = 16
    ys, zs = np.meshgrid(np.arange(n + 1) - 8.5, np.arange(n + 1) - 8.5, indexing="ij")
    pos = np.stack([np.zeros_like(ys), ys, zs], -1).reshape(-1, 3)
    quads = [(i * (n + 1) + j, i * (n + 1) + j + 1, (i + 1) * (n + 1) + j + 1, (i + 1) * (n + 1) + j)
             for i in range(n) for j in range(n)]
    idx = [(a, b, c) for a, b, c, d in quads] + [(a, c, d) for a, b, c, d in quads]
    return Mesh(pos, np.array(idx), np.zeros(len(idx), np.uint16), np.array([1], np.uint8))


def overlap_cases():
    ico = convex()
    dup = Mesh(ico.positions, np.concatenate([ico.indices, ico.indices]))
    two = Mesh(ico.positions, ico.indices, np.zeros(ico.triangles, np.uint16), np.array([1], np.uint8))
    return {"convex": ico, "duplicated": dup, "two_sided": two, "noisy": meshes.icosphere(3, 0.05, 2),
            "heightfield": meshes.heightfield(40, 0.1, 2), "torus": meshes.torus(1.0, 0.3, 40, 16, 0.01, 3),
            "tiles": meshes.four_tiles()}


def test_overlap_closed_forms():
    o = Oracle("oracle")
    cases = overlap_cases()
    c, r = o.bounds(cases["convex"])
    for screen, ortho, persp in ((40, 8, 0), (63.5, 4, 4), (128, 3, 3)):
        assert o.overlap(cases["convex"], c, r, screen, ortho, persp, 7) == 1.0
        assert o.overlap(cases["duplicated"], c, r, screen, ortho, persp, 7) == 2.0
        assert o.overlap(cases["two_sided"], c, r, screen, ortho, persp, 7) == 2.0
    assert o.overlap(cases["convex"], c, r * 0.5, 40, 4, 0, 7) == float("inf")
    # Pixel centres on shared edges: each is counted by exactly one face.
    assert o.overlap(pixel_grid(), np.zeros(3), 28.0, 56, 1, 0, 0) == 1.0
    assert o.overlap(cases["noisy"], *o.bounds(cases["noisy"]), 64, 6, 2, 7) > 1.0


def test_overlap_planted_mistake_is_caught(tmp_path):
    import shutil
    import subprocess
    from pathlib import Path
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    root = Path(__file__).resolve().parent.parent
    out = tmp_path / "liboracle_plant7.so"
    subprocess.run([gxx, "-std=c++20", "-O2", "-fPIC", "-ffp-contract=off", "-shared", "-DORACLE_PLANT=7",
                    "-o", str(out), str(root / "oracle" / "audit_oracle.cpp")], check=True)
    planted = Oracle("oracle", out)
    assert planted.overlap(pixel_grid(), np.zeros(3), 28.0, 56, 1, 0, 0) > 1.0


@pytest.mark.gpu
def test_gpu_overlap_matches_oracle():
    from blitzremesher_amd import Audit
    o = Oracle("oracle")
    with Audit(0) as audit:
        for name, mesh in overlap_cases().items():
            c, r = o.bounds(mesh)
            for screen, ortho, persp, seed in ((40, 8, 0, 7), (63.5, 4, 4, 11), (200, 3, 2, 0xB1172026)):
                got = audit.overlap(mesh, c, r, screen, ortho, persp, seed)
                want = o.overlap(mesh, c, r, screen, ortho, persp, seed)
                assert got == want, (name, screen, got, want)
        ico = convex()
        c, r = o.bounds(ico)
        assert audit.overlap(ico, c, r, 40, 8, 0, 7) == 1.0
        assert audit.overlap(pixel_grid(), np.zeros(3), 28.0, 56, 1, 0, 0) == 1.0
        assert audit.overlap(ico, c, r * 0.5, 40, 4, 0, 7) == float("inf")
        with pytest.raises(Exception):
            audit.overlap(ico, c, r, 40, 0, 0, 7)
