"""Collapse actions: topology, legality, independent batches, trial edits
(src/neural/action.cpp ActionState; GPU counterpart
src/neural/action_gpu.cu:178-343,487-643,757-808,865-899).

CPU: the oracle's restatement equals golden fixtures written by the
reference's own ActionState and, where that build is present, the build
itself; and it keeps what a legal collapse must keep: a closed manifold stays
closed and manifold with the same Euler characteristic after any batch, each
interior collapse removes two faces and each boundary collapse one, and a
tetrahedron has no legal collapse. GPU: device topology, legality, selection
and trials equal the golden fixtures and the oracle exactly.
"""
from __future__ import annotations

import hashlib
import json
from functools import lru_cache
from pathlib import Path

import numpy as np
import pytest

from blitzremesher_amd import Mesh, meshes
from action_meshes import action_cases
from oracle_lib import Actions, available

GOLDEN = Path(__file__).resolve().parent / "golden"
needs_ref = pytest.mark.skipif(not available("ref"), reason="reference CPU build (oracle/_ref) not present")


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.uint32).tobytes()).hexdigest()


@lru_cache(maxsize=1)
def golden():
    return json.loads((GOLDEN / "action_golden.json").read_text())


@lru_cache(maxsize=1)
def golden_meshes():
    data = np.load(GOLDEN / "action_meshes.npz")
    out = {}
    for n in sorted({k.split("/")[0] for k in data.files}):
        mats = data[f"{n}/materials"] if f"{n}/materials" in data.files else None
        uv = data[f"{n}/uv"] if f"{n}/uv" in data.files else None
        out[n] = (Mesh(data[f"{n}/positions"], data[f"{n}/indices"], mats), uv)
    return out


def check_against_golden(make_state):
    ms = golden_meshes()
    for case in golden()["cases"]:
        mesh, uv = ms[case["mesh"]]
        st = make_state(mesh, uv)
        assert (len(st.legal), sha(st.legal)) == (case["legal"], case["legal_sha256"]), case["mesh"]
        for b in case["batches"]:
            order = np.random.default_rng(b["seed"]).permutation(len(st.legal)).astype(np.uint32)
            picked = st.independent(order, b["maximum"], b["maximum_removed"])
            assert picked.tolist() == b["selected"], (case["mesh"], b["seed"])
            if len(picked):
                t = st.trial(picked)
                assert (len(t), sha(t)) == (b["trial_faces"], b["trial_sha256"]), (case["mesh"], b["seed"])


def test_oracle_matches_golden():
    check_against_golden(lambda m, uv: Actions("oracle", m, uv))


@needs_ref
@pytest.mark.parametrize("seed", [21, 22])
def test_oracle_matches_reference_build(seed):
    cases = {"icosphere": (meshes.icosphere(3, 0.03, seed), None),
             "heightfield": (meshes.jitter(meshes.heightfield(18, 0.4, seed), 0.01, seed, 0.5), None),
             "torus": (meshes.torus(1.0, 0.3, 20, 9, 0.02, seed), None)}
    cases.update(action_cases())
    for name, (mesh, uv) in cases.items():
        a, r = Actions("oracle", mesh, uv), Actions("ref", mesh, uv)
        assert np.array_equal(a.legal, r.legal), name
        rng = np.random.default_rng(seed)
        for maximum, budget in ((64, 10 ** 9), (3, 10 ** 9), (64, 5), (64, 1)):
            order = rng.permutation(len(a.legal)).astype(np.uint32)
            s1, s2 = a.independent(order, maximum, budget), r.independent(order, maximum, budget)
            assert np.array_equal(s1, s2), (name, maximum, budget)
            if len(s1):
                assert np.array_equal(a.trial(s1), r.trial(s2)), (name, maximum, budget)


def edge_multiplicity(indices: np.ndarray, geometry: np.ndarray) -> np.ndarray:
    g = geometry[indices]
    e = np.sort(np.concatenate([g[:, [0, 1]], g[:, [1, 2]], g[:, [2, 0]]]), 1)
    _, counts = np.unique(e, axis=0, return_counts=True)
    return counts


def euler(indices: np.ndarray) -> int:
    e = np.sort(np.concatenate([indices[:, [0, 1]], indices[:, [1, 2]], indices[:, [2, 0]]]), 1)
    return len(np.unique(indices)) - len(np.unique(e, axis=0)) + len(indices)


@pytest.mark.parametrize("name", ["icosphere", "torus"])
def test_legal_batches_keep_closed_manifolds(name):
    mesh, _ = action_cases()[name]
    st = Actions("oracle", mesh)
    geometry = np.arange(len(mesh.positions))
    chi = euler(mesh.indices)
    rng = np.random.default_rng(4)
    for _ in range(5):
        order = rng.permutation(len(st.legal)).astype(np.uint32)
        picked = st.independent(order, 64, 10 ** 9)
        t = st.trial(picked)
        assert len(t) == mesh.triangles - 2 * len(picked)
        assert np.all(edge_multiplicity(t, geometry) == 2)
        assert euler(t) == chi
    for k in range(0, len(st.legal), max(1, len(st.legal) // 40)):
        assert len(st.trial(np.array([k]))) == mesh.triangles - 2


def test_boundary_collapses_remove_one_face():
    mesh = meshes.grid(6)
    st = Actions("oracle", mesh)
    n = int(round(np.sqrt(len(mesh.positions))))  # vertices per side
    on_boundary = lambda v: v // n in (0, n - 1) or v % n in (0, n - 1)  # noqa: E731
    kinds = set()
    for k, (u, v) in enumerate(st.legal):
        removed = mesh.triangles - len(st.trial(np.array([k])))
        if on_boundary(int(u)):
            assert on_boundary(int(v)) and removed == 1
        kinds.add(removed)
    assert kinds == {1, 2}


def test_tetrahedron_has_no_legal_collapse():
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    idx = np.array([[0, 2, 1], [0, 1, 3], [1, 2, 3], [0, 3, 2]], np.uint32)
    assert len(Actions("oracle", Mesh(pos, idx)).legal) == 0


# ----------------------------------------------------------------------- GPU

def face_multiset(idx: np.ndarray):
    return sorted(map(tuple, np.asarray(idx).reshape(-1, 3).tolist()))


@pytest.mark.gpu
def test_gpu_actions_match_golden():
    from blitzremesher_amd.actions import ActionSet
    check_against_golden(lambda m, uv: ActionSet(m, uv))


@pytest.mark.gpu
def test_gpu_actions_match_oracle():
    from blitzremesher_amd.actions import ActionSet
    cases = {"icosphere4": (meshes.icosphere(4, 0.03, 7), None),
             "heightfield": (meshes.jitter(meshes.heightfield(64, 0.5, 8), 0.01, 8, 0.5), None),
             "torus": (meshes.torus(1.0, 0.3, 60, 24, 0.02, 9), None)}
    cases.update(action_cases())
    for name, (mesh, uv) in cases.items():
        gpu, cpu = ActionSet(mesh, uv), Actions("oracle", mesh, uv)
        assert np.array_equal(gpu.legal, cpu.legal), name
        rng = np.random.default_rng(31)
        for maximum, budget in ((64, 10 ** 9), (7, 10 ** 9), (64, 9), (64, 1)):
            order = rng.permutation(len(cpu.legal)).astype(np.uint32)
            s1, s2 = gpu.independent(order, maximum, budget), cpu.independent(order, maximum, budget)
            assert np.array_equal(s1, s2), (name, maximum, budget)
            if len(s1):
                assert np.array_equal(gpu.trial(s1), cpu.trial(s2)), (name, maximum, budget)
        gpu.close()


@pytest.mark.gpu
def test_gpu_action_patches_equal_trials():
    from blitzremesher_amd.actions import ActionSet
    for name in ("heightfield", "seam", "junction", "torus"):
        mesh, uv = action_cases()[name]
        with ActionSet(mesh, uv) as gpu:
            rows = np.arange(0, len(gpu.legal), max(1, len(gpu.legal) // 60), dtype=np.uint32)
            p = gpu.patches(rows)
            assert len(p) == len(rows)
            for i, r in enumerate(rows):
                assert np.all(np.diff(p.removed[p.removed_offsets[i]:p.removed_offsets[i + 1]].astype(np.int64)) > 0)
                assert face_multiset(p.candidate(mesh, i).indices) == face_multiset(gpu.trial(np.array([r]))), (name, r)


@pytest.mark.gpu
def test_gpu_trial_rejects_overlapping_actions():
    from blitzremesher_amd._lib import BlitzError
    from blitzremesher_amd.actions import ActionSet
    mesh = meshes.icosphere(2, 0.0, 1)
    with ActionSet(mesh) as gpu:
        u = gpu.legal[0, 0]
        same = np.nonzero(gpu.legal[:, 0] == u)[0][:2].astype(np.uint32)
        with pytest.raises(BlitzError):
            gpu.trial(same)
        with pytest.raises(BlitzError):
            gpu.independent(np.array([len(gpu.legal)], np.uint32), 4, 10)


@pytest.mark.gpu
def test_gpu_action_patches_audit_like_full_candidates():
    # Every selected collapse of a seeded rank order, audited as a patch of the
    # bound source, equals the full audit of its trial mesh.
    from blitzremesher_amd import Audit, EvalSettings
    from blitzremesher_amd.actions import ActionSet
    from blitzremesher_amd.audit import bounds
    mesh = meshes.jitter(meshes.heightfield(48, 0.2, 4), 0.004, 4, 0.5)
    # limit 0.75 at ss 4 passes a view only with zero coverage distance
    # (allowance 2*sqrt(2)/4 + 1e-6): both verdicts occur.
    s = EvalSettings(profile=0, orthographic=12, perspective=0, screen_size=40, supersample=4, max_supersample=4,
                     limit=0.75, force_two_sided=True)
    center, radius = bounds(mesh)
    with ActionSet(mesh) as gpu, Audit(0) as audit:
        rows = np.random.default_rng(2).choice(len(gpu.legal), 96, replace=False).astype(np.uint32)
        p = gpu.patches(rows)
        binding = audit.bind(mesh, center, radius, s)
        got, _ = audit.evaluate_patches(binding, p, s)
        ref = audit.upload(mesh)
        for i, r in enumerate(rows):
            cand = audit.upload(Mesh(mesh.positions, gpu.trial(np.array([r]))))
            assert got[i] == audit.evaluate(ref, cand, center, radius, s), (i, r)
            cand.release()
        assert any(not m.passed for m in got) and any(m.passed for m in got)
