"""The device classical reducer (blitz::reduce with endpoint output,
src/simplify.cpp:43-329) against the reference's own reduce(): the golden
reductions written from the reference build (icosphere to 900 and 300
triangles, torus, heightfield to 75 %), and the reference build itself on
closed and open meshes at several budgets (the last pass's budget cut and
its one-face tail included), with material junctions and vertex normals (the
bend cost), on the raster-audit configuration's 512 x 512 heightfield and on
the LOD-chain icosphere. Equal means the same triangles in the same order."""
from __future__ import annotations

import numpy as np
import pytest

from blitzremesher_amd import Mesh, meshes
from blitzremesher_amd.audit import reduce_reuse_device
from golden_data import golden_meshes
from oracle_lib import Oracle, available

GOLDEN = [("ico3", "ico3_qem_900", 900), ("ico3", "ico3_qem_300", 300), ("torus60", "torus60_qem_1500", 1500),
          ("heightfield48", "heightfield48_qem", None)]


@pytest.mark.gpu
@pytest.mark.parametrize("source,reduced,target", GOLDEN)
def test_gpu_reduce_matches_golden(source, reduced, target):
    g = golden_meshes()
    mesh = g[source]
    target = int(mesh.triangles * 0.75) if target is None else target
    report = {}
    got = reduce_reuse_device(mesh, target, report=report)
    print(source, target, report, flush=True)
    assert np.array_equal(got.indices, g[reduced].indices)
    assert report["passes"] >= 1 and report["collapses"] > 0


def with_materials(mesh: Mesh, seed: int) -> Mesh:
    """Two spatial material regions (by the sign of the face centroid's x)
    and a few scattered faces of a third material."""
    rng = np.random.default_rng(seed)
    centroid_x = mesh.positions[mesh.indices].mean(axis=1)[:, 0]
    mat = (centroid_x > np.median(centroid_x)).astype(np.uint16)
    mat[rng.random(mesh.triangles) < 0.003] = 2
    return Mesh(mesh.positions, mesh.indices, mat)


def with_normals(mesh: Mesh) -> Mesh:
    p = mesh.positions.astype(np.float64)
    n = np.zeros_like(p)
    f = mesh.indices.astype(np.int64)
    fn = np.cross(p[f[:, 1]] - p[f[:, 0]], p[f[:, 2]] - p[f[:, 0]])
    for k in range(3):
        np.add.at(n, f[:, k], fn)
    return Mesh(mesh.positions, mesh.indices, normals=n.astype(np.float32))


def live_cases():
    ico4 = meshes.icosphere(4, 0.03, 2)
    hf = meshes.heightfield(64, 0.1, 2)
    box = meshes.extruded_box(12, 0.5, 7)
    torus = meshes.torus(1.0, 0.3, 48, 20, 0.01, 3)
    return [
        ("ico4_half", ico4, ico4.triangles // 2),
        ("ico4_tenth", ico4, ico4.triangles // 10),
        ("ico4_tiny", ico4, 5),
        ("heightfield64_odd", hf, int(hf.triangles * 0.6) + 1),
        ("heightfield64_quarter", hf, hf.triangles // 4),
        ("box_half", box, box.triangles // 2),
        ("torus_third", torus, torus.triangles // 3),
        ("ico4_materials", with_materials(ico4, 5), ico4.triangles // 3),
        ("heightfield64_materials", with_materials(hf, 6), hf.triangles // 2 + 1),
        ("torus_normals", with_normals(torus), torus.triangles // 3),
        ("unreduced", ico4, ico4.triangles + 10),
    ]


@pytest.mark.gpu
@pytest.mark.skipif(not available("ref"), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("case", range(len(live_cases())))
def test_gpu_reduce_matches_reference_build(case):
    name, mesh, target = live_cases()[case]
    ref = Oracle("ref")
    want = ref.reduce_reuse_streams(mesh, target)
    report = {}
    got = reduce_reuse_device(mesh, target, report=report)
    print(name, mesh.triangles, "->", len(want), report, flush=True)
    assert np.array_equal(got.indices, want), name


@pytest.mark.gpu
@pytest.mark.skipif(not available("ref"), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("which", ["heightfield512", "icosphere5"])
def test_gpu_reduce_matches_reference_build_at_config_size(which):
    if which == "heightfield512":
        mesh = meshes.heightfield(512, 0.1, 2)
        target = int(round(mesh.triangles * 0.75))
    else:
        mesh = meshes.icosphere(5, 0.02, 1)
        target = mesh.triangles // 4
    want = Oracle("ref").reduce_reuse(mesh, target).indices
    report = {}
    got = reduce_reuse_device(mesh, target, report=report)
    print(which, mesh.triangles, "->", len(want), report, flush=True)
    assert np.array_equal(got.indices, want)
