"""Local patches and the bound-reference patch audit.

CPU: the patch generator's contract (seeded, region sizes in range, ascending
region ids, added faces over region vertices, edge-manifold candidates) and
the candidate construction. GPU: every patch measurement equals, field for
field, the full audit of that patch's candidate mesh (blitz_hip_audit_evaluate)
and, on a sample, the CPU oracle; cases cover refinement, first-failure stops,
perspective views, empty and identical patches, holes and rejected inputs.
"""
from __future__ import annotations

import numpy as np
import pytest

from blitzremesher_amd import Audit, EvalSettings, Mesh, meshes
from blitzremesher_amd._lib import BlitzError
from blitzremesher_amd.audit import Patches, bounds, local_patches
from oracle_lib import Oracle


def manifold_edges(indices: np.ndarray) -> int:
    e = np.sort(np.concatenate([indices[:, [0, 1]], indices[:, [1, 2]], indices[:, [2, 0]]]), 1)
    _, counts = np.unique(e, axis=0, return_counts=True)
    return int(counts.max())


def test_local_patches_contract():
    mesh = meshes.heightfield(48, 0.1, 2)
    p = local_patches(mesh, 64, 11, 16, 200)
    q = local_patches(mesh, 64, 11, 16, 200)
    assert all(np.array_equal(a, b) for a, b in zip(
        (p.removed_offsets, p.removed, p.added_offsets, p.added), (q.removed_offsets, q.removed, q.added_offsets,
                                                                    q.added)))
    sizes = np.diff(p.removed_offsets)
    assert sizes.min() >= 16 and sizes.max() <= 200 and len(p) == 64
    collapsed = 0
    for i in range(len(p)):
        r = p.removed[p.removed_offsets[i]:p.removed_offsets[i + 1]]
        a = p.added[p.added_offsets[i]:p.added_offsets[i + 1]]
        assert np.all(np.diff(r.astype(np.int64)) > 0)
        assert set(a.reshape(-1).tolist()) <= set(mesh.indices[r].reshape(-1).tolist())
        assert np.all((a[:, 0] != a[:, 1]) & (a[:, 1] != a[:, 2]) & (a[:, 0] != a[:, 2]))
        c = p.candidate(mesh, i)
        assert c.triangles == mesh.triangles - r.size + a.shape[0]
        assert manifold_edges(c.indices) <= 2
        collapsed += r.size - a.shape[0]
    assert collapsed > 0


def test_patch_candidate_order():
    mesh = meshes.grid(3)
    p = Patches(np.array([0, 2], np.uint32), np.array([1, 4], np.uint32), np.array([0, 1], np.uint32),
                np.array([[0, 5, 10]], np.uint32))
    c = p.candidate(mesh, 0)
    keep = [f for f in range(mesh.triangles) if f not in (1, 4)]
    assert np.array_equal(c.indices[:-1], mesh.indices[keep])
    assert np.array_equal(c.indices[-1], [0, 5, 10])


# ----------------------------------------------------------------------- GPU

def full_audit(audit, mesh, patches, center, radius, s):
    ref = audit.upload(mesh)
    out = []
    for i in range(len(patches)):
        cand = audit.upload(patches.candidate(mesh, i))
        out.append(audit.evaluate(ref, cand, center, radius, s))
        cand.release()
    ref.release()
    return out


def check_patches(mesh, patches, s, expect_fallbacks=None, oracle_sample=0):
    center, radius = bounds(mesh)
    with Audit(0) as audit:
        binding = audit.bind(mesh, center, radius, s)
        got, fallbacks = audit.evaluate_patches(binding, patches, s)
        want = full_audit(audit, mesh, patches, center, radius, s)
    for i, (g, w) in enumerate(zip(got, want)):
        assert g.as_tuple() == w.as_tuple(), f"patch {i}\n patch={g}\n full={w}"
    if expect_fallbacks is not None:
        assert fallbacks == expect_fallbacks
    if oracle_sample:
        oracle = Oracle("oracle")
        for i in range(0, len(patches), max(1, len(patches) // oracle_sample)):
            cpu = oracle.evaluate(mesh, patches.candidate(mesh, i), center, radius, s)
            assert got[i].as_tuple() == cpu.as_tuple(), f"patch {i} vs oracle"
    return got, fallbacks


@pytest.mark.gpu
def test_gpu_patches_match_full_audit_heightfield():
    mesh = meshes.heightfield(96, 0.1, 2)
    p = local_patches(mesh, 48, 3, 16, 200)
    s = EvalSettings(profile=0, orthographic=24, perspective=0, screen_size=40, supersample=4, max_supersample=4,
                     limit=1e3, force_two_sided=True)
    got, fallbacks = check_patches(mesh, p, s, expect_fallbacks=0, oracle_sample=4)
    assert any(m.coverage > 0 for m in got)


@pytest.mark.gpu
def test_gpu_patches_first_failure_and_refinement():
    # A tight limit makes views fail (stop at the first failure) and leaves
    # some uncertain (refinement to ss 8, audited by the full path).
    mesh = meshes.icosphere(3, 0.02, 1)
    p = local_patches(mesh, 40, 5, 8, 60)
    s = EvalSettings(profile=0, orthographic=12, perspective=6, screen_size=48, supersample=4, max_supersample=8,
                     limit=0.75, max_changed_area=1.0)
    got, fallbacks = check_patches(mesh, p, s, oracle_sample=3)
    assert any(not m.passed for m in got)


@pytest.mark.gpu
def test_gpu_patches_perspective_and_one_sided():
    mesh = meshes.torus(1.0, 0.3, 40, 16, 0.01, 3)
    p = local_patches(mesh, 32, 9, 16, 120)
    s = EvalSettings(profile=0, orthographic=8, perspective=8, screen_size=64, supersample=2, max_supersample=2,
                     limit=1e3)
    check_patches(mesh, p, s, expect_fallbacks=0, oracle_sample=3)


@pytest.mark.gpu
def test_gpu_patch_edge_cases():
    mesh = meshes.heightfield(32, 0.1, 2)
    n = mesh.triangles
    tail = mesh.indices[n - 3:]
    cases = [
        (np.zeros(0, np.uint32), np.zeros((0, 3), np.uint32)),                       # empty: identical
        (np.arange(n - 3, n, dtype=np.uint32), tail.copy()),                         # trailing faces re-added
        (np.array([5, 6, 7, 200], np.uint32), np.zeros((0, 3), np.uint32)),          # hole
        (np.zeros(0, np.uint32), mesh.indices[10:12].copy()),                        # duplicated faces appended
        (np.array([0], np.uint32), mesh.indices[[0]][:, ::-1].copy()),               # flipped face
    ]
    ro, rv, ao, av = [0], [], [0], []
    for r, a in cases:
        rv.append(r)
        av.append(a)
        ro.append(ro[-1] + r.size)
        ao.append(ao[-1] + a.shape[0])
    p = Patches(np.array(ro, np.uint32), np.concatenate(rv).astype(np.uint32), np.array(ao, np.uint32),
                np.concatenate(av).astype(np.uint32))
    s = EvalSettings(profile=0, orthographic=10, perspective=2, screen_size=30, supersample=4, max_supersample=4,
                     limit=1e3)
    got, _ = check_patches(mesh, p, s, oracle_sample=5)
    assert got[0].views_evaluated == 0 and got[1].views_evaluated == 0
    assert got[2].changed_area > 0


@pytest.mark.gpu
def test_gpu_patches_full_size_views():
    # The raster-audit workload's view set and raster (64 views at 1024^2)
    # on a smaller heightfield, patches against full audits.
    mesh = meshes.heightfield(160, 0.1, 2)
    p = local_patches(mesh, 24, 21, 16, 200)
    s = EvalSettings(profile=0, orthographic=64, perspective=0, rotation_seed=0xB1172026, screen_size=120,
                     supersample=8, max_supersample=8, limit=1e3, force_two_sided=True)
    check_patches(mesh, p, s, expect_fallbacks=0)


@pytest.mark.gpu
def test_gpu_patch_inputs_rejected():
    mesh = meshes.heightfield(16, 0.1, 2)
    center, radius = bounds(mesh)
    s = EvalSettings(profile=0, orthographic=4, perspective=0, screen_size=20, supersample=2, max_supersample=2)
    with Audit(0) as audit:
        b = audit.bind(mesh, center, radius, s)
        bad_order = Patches(np.array([0, 2], np.uint32), np.array([7, 3], np.uint32), np.array([0, 0], np.uint32),
                            np.zeros((0, 3), np.uint32))
        with pytest.raises(BlitzError):
            audit.evaluate_patches(b, bad_order, s)
        bad_vertex = Patches(np.array([0, 0], np.uint32), np.zeros(0, np.uint32), np.array([0, 1], np.uint32),
                             np.array([[0, 1, 10 ** 6]], np.uint32))
        with pytest.raises(BlitzError):
            audit.evaluate_patches(b, bad_vertex, s)
        other = EvalSettings(**{**s.__dict__, "supersample": 4, "max_supersample": 4})
        ok = Patches(np.array([0, 1], np.uint32), np.array([0], np.uint32), np.array([0, 0], np.uint32),
                     np.zeros((0, 3), np.uint32))
        with pytest.raises(BlitzError):
            audit.evaluate_patches(b, ok, other)
        with_materials = Mesh(mesh.positions, mesh.indices, np.zeros(mesh.triangles, np.uint16),
                              np.array([1], np.uint8))
        b2 = audit.bind(with_materials, center, radius, s)
        with pytest.raises(BlitzError):
            audit.evaluate_patches(b2, ok, s)
