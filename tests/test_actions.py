"""Collapse actions: topology, legality, independent batches, trial edits
(src/neural/action.cpp ActionState; GPU counterpart
src/neural/action_gpu.cu:178-343,487-643,757-808,865-899).

CPU: the oracle's restatement equals golden fixtures written by the
reference's own ActionState and, where that build is present, the build
itself; and it keeps what a legal collapse must keep: a closed manifold stays
closed and manifold with the same Euler characteristic after any batch, each
interior collapse removes two faces and each boundary collapse one, and a
tetrahedron has no legal collapse. The oracle's restatement of
execute_actions (rank, greedy batch, trial, gate, commit or halve) equals the
reference's own loop with its CPU audit as the gate, output and work counters.
GPU: device topology, legality, ranking, selection, trials and the executor
(with the HIP audit as the gate) equal the golden fixtures and the oracle
exactly.
"""
from __future__ import annotations

import hashlib
import json
from functools import lru_cache
from pathlib import Path

import numpy as np
import pytest

from blitzremesher_amd import EvalSettings, Mesh, meshes
from action_meshes import action_cases, executor_cases
from oracle_lib import Actions, Oracle, available, execute_actions, ref_execute_actions

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
        opt = {k: data[f"{n}/{k}"] if f"{n}/{k}" in data.files else None for k in ("materials", "double_sided", "uv")}
        out[n] = (Mesh(data[f"{n}/positions"], data[f"{n}/indices"], opt["materials"], opt["double_sided"]),
                  opt["uv"])
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


def oracle_gate(mesh, center, radius, settings):
    o = Oracle("oracle")
    return lambda idx: o.evaluate(mesh, Mesh(mesh.positions, idx, double_sided=mesh.double_sided), center, radius,
                                  settings).passed


def test_oracle_executor_matches_golden():
    ms = golden_meshes()
    runs = golden()["executions"]
    assert len(runs) >= 40 and any(r["stats"]["rejected"] for r in runs)
    for r in runs:
        mesh, _ = ms[r["mesh"]]
        gate = EvalSettings(**r["gate"])
        out, stats = execute_actions(mesh, r["target"], r["budget"], r["ranking"], r["batch"],
                                     oracle_gate(mesh, r["center"], r["radius"], gate))
        assert (len(out), sha(out), stats) == (r["faces"], r["indices_sha256"], r["stats"]), r


@needs_ref
def test_oracle_executor_matches_reference_build():
    gate = EvalSettings(profile=0, screen_size=20, limit=1.5, supersample=2, max_supersample=8, orthographic=3,
                        perspective=2, rotation_seed=77)
    o = Oracle("oracle")
    for mesh in (meshes.icosphere(2, 0.05, 9), meshes.jitter(meshes.heightfield(10, 0.4, 9), 0.01, 9, 0.5),
                 meshes.torus(1.0, 0.3, 10, 6, 0.03, 9)):
        center, radius = o.bounds(mesh)
        for ranking, batch in ((1, 2), (3, 8), (4, 64)):
            a = execute_actions(mesh, mesh.triangles // 4, 30, ranking, batch, oracle_gate(mesh, center, radius, gate))
            b = ref_execute_actions(mesh, center, radius, gate, mesh.triangles // 4, 30, ranking, batch)
            assert np.array_equal(a[0], b[0]) and a[1] == b[1], (ranking, batch, a[1], b[1])


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


def bipyramid(k: int) -> Mesh:
    """A closed k-gon bipyramid: two apexes of valence k."""
    t = 2 * np.pi * np.arange(k) / k
    p = np.zeros((k + 2, 3), np.float32)
    p[:k, 0], p[:k, 1] = np.cos(t), np.sin(t)
    p[k, 2], p[k + 1, 2] = 1.0, -1.0
    f = [(k, i, (i + 1) % k) for i in range(k)] + [(k + 1, (i + 1) % k, i) for i in range(k)]
    return Mesh(p, np.array(f, np.uint32))


def bowtie(k: int) -> Mesh:
    """Two closed cones joined at their apex (vertex 0): the apex's link is
    two separate cycles of k vertices, so it is not one fan."""
    t = 2 * np.pi * np.arange(k) / k
    ring = np.stack([np.cos(t), np.sin(t), np.zeros(k)], 1)
    p = np.concatenate([[[0, 0, 0]], ring + [0, 0, 1], [[0, 0, 1.2]], ring + [0, 0, -1], [[0, 0, -1.2]]])
    up, cap_up, down, cap_down = 1, 1 + k, 2 + k, 2 + 2 * k
    f = [(0, up + (i + 1) % k, up + i) for i in range(k)] + [(cap_up, up + i, up + (i + 1) % k) for i in range(k)]
    f += [(0, down + i, down + (i + 1) % k) for i in range(k)]
    f += [(cap_down, down + (i + 1) % k, down + i) for i in range(k)]
    return Mesh(p.astype(np.float32), np.array(f, np.uint32))


@pytest.mark.gpu
def test_gpu_actions_match_oracle():
    """Includes vertices whose lists take the wave-per-vertex sort and fan
    test (latitude-longitude sphere poles, a 300-valence bipyramid, a
    non-manifold apex of two 80-face cones) and ones past the LDS lists
    (a 1,100-valence bipyramid)."""
    from blitzremesher_amd.actions import ActionSet
    cases = {"icosphere4": (meshes.icosphere(4, 0.03, 7), None),
             "heightfield": (meshes.jitter(meshes.heightfield(64, 0.5, 8), 0.01, 8, 0.5), None),
             "torus": (meshes.torus(1.0, 0.3, 60, 24, 0.02, 9), None),
             "latlong_sphere": (meshes.training_mesh("sphere", 25000, 17), None),
             "bipyramid300": (bipyramid(300), None), "bowtie80": (bowtie(80), None),
             "bipyramid1100": (bipyramid(1100), None)}
    cases.update(action_cases())
    for name, (mesh, uv) in cases.items():
        gpu, cpu = ActionSet(mesh, uv), Actions("oracle", mesh, uv)
        print(name, len(cpu.legal), flush=True)
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


@pytest.mark.gpu
def test_gpu_rank_matches_oracle():
    from blitzremesher_amd.actions import ActionSet
    o = Oracle("oracle")
    for name in ("icosphere", "heightfield", "folded", "seam", "torus"):
        mesh, uv = action_cases()[name]
        _, radius = o.bounds(mesh)
        cpu = Actions("oracle", mesh, uv)
        with ActionSet(mesh, uv) as gpu:
            for ranking in (1, 3, 4):
                assert np.array_equal(gpu.rank(ranking), cpu.rank(ranking, max(1e-20, 2 * radius))), (name, ranking)


@pytest.mark.gpu
def test_gpu_learned_rank_orders_scorer_logits():
    from blitzremesher_amd.actions import ActionSet
    from blitzremesher_amd.scorer import BF16, Scorer, he_weights
    mesh = meshes.torus(1.0, 0.3, 40, 16, 0.01, 3)
    o = Oracle("oracle")
    center, radius = o.bounds(mesh)
    with Scorer(he_weights(4)) as scorer, ActionSet(mesh) as gpu:
        g = scorer.graph(mesh, center, radius)
        logits = scorer.forward(g, BF16)[:, 0]
        pairs = g.actions_array()
        key = {(int(a), int(b)): i for i, (a, b) in enumerate(pairs)}
        scores = np.array([logits[key[(int(u), int(v))]] for u, v in gpu.legal], np.float32)
        want = np.argsort(-(scores + np.float32(0)), kind="stable").astype(np.uint32)
        assert np.array_equal(gpu.rank(0, scorer, g), want)


@pytest.mark.gpu
@pytest.mark.parametrize("cache", [False, True])
def test_gpu_executor_matches_golden(cache):
    from blitzremesher_amd import Audit
    from blitzremesher_amd.actions import execute_actions as gpu_execute
    ms = golden_meshes()
    with Audit(0, cache_rasters=cache) as audit:
        for r in golden()["executions"]:
            mesh, _ = ms[r["mesh"]]
            out, stats = gpu_execute(audit, mesh, r["center"], r["radius"], EvalSettings(**r["gate"]), r["target"],
                                     r["budget"], r["ranking"], r["batch"])
            got = {k: stats[k] for k in ("ranked", "trials", "accepted", "rejected")}
            assert (len(out), sha(out), got) == (r["faces"], r["indices_sha256"], r["stats"]), (r, got)


@pytest.mark.gpu
@pytest.mark.parametrize("cache", [False, True])
def test_gpu_executor_matches_oracle(cache):
    from blitzremesher_amd import Audit
    from blitzremesher_amd.actions import execute_actions as gpu_execute
    gate = EvalSettings(profile=0, screen_size=32, limit=1.25, supersample=2, max_supersample=8, orthographic=4,
                        perspective=2, rotation_seed=5)
    o = Oracle("oracle")
    with Audit(0, cache_rasters=cache) as audit:
        for mesh in (meshes.icosphere(3, 0.03, 11), meshes.jitter(meshes.heightfield(16, 0.3, 11), 0.01, 11, 0.5)):
            center, radius = o.bounds(mesh)
            for ranking, batch in ((3, 16), (4, 64), (1, 4)):
                want = execute_actions(mesh, mesh.triangles // 3, 50, ranking, batch,
                                       oracle_gate(mesh, center, radius, gate))
                out, stats = gpu_execute(audit, mesh, center, radius, gate, mesh.triangles // 3, 50, ranking, batch)
                got = {k: stats[k] for k in ("ranked", "trials", "accepted", "rejected")}
                assert np.array_equal(out, want[0]) and got == want[1], (ranking, batch, got, want[1])


@pytest.mark.gpu
@pytest.mark.parametrize("cache", [False, True])
def test_gpu_executor_conservative_gate_matches_oracle(cache):
    # A conservative screen fails a view only on a proved coverage violation
    # (coverage - 2*sqrt(2)/ss - 1e-6 > limit) and never refines; the gate's
    # verdict-only Hausdorff search must stop on that rule, not the full one.
    # supersample == max_supersample puts every view on the last rung.
    from blitzremesher_amd import Audit
    from blitzremesher_amd.actions import execute_actions as gpu_execute
    o = Oracle("oracle")
    with Audit(0, cache_rasters=cache) as audit:
        for mesh, ss, limit in ((meshes.icosphere(3, 0.03, 11), 4, 0.75), (meshes.icosphere(3, 0.03, 13), 4, 0.0),
                                (meshes.jitter(meshes.heightfield(16, 0.3, 11), 0.01, 11, 0.5), 4, 0.2)):
            gate = EvalSettings(profile=0, screen_size=32, limit=limit, supersample=ss, max_supersample=ss,
                                orthographic=4, perspective=2, rotation_seed=5, conservative_screen=True)
            center, radius = o.bounds(mesh)
            for ranking, batch in ((4, 64), (3, 8)):
                want = execute_actions(mesh, mesh.triangles // 4, 40, ranking, batch,
                                       oracle_gate(mesh, center, radius, gate))
                out, stats = gpu_execute(audit, mesh, center, radius, gate, mesh.triangles // 4, 40, ranking, batch)
                got = {k: stats[k] for k in ("ranked", "trials", "accepted", "rejected")}
                assert np.array_equal(out, want[0]) and got == want[1], (ss, limit, ranking, batch, got, want[1])


@pytest.mark.gpu
def test_gpu_learned_executor_matches_oracle_loop():
    # The executor's learned order: per mesh state the scorer's BF16 logits of
    # that state's graph. The oracle loop is driven by the same logits (read
    # back from the scorer per state), so output and counters must agree.
    from blitzremesher_amd import Audit
    from blitzremesher_amd.actions import execute_actions as gpu_execute
    from blitzremesher_amd.scorer import BF16, Scorer, he_weights
    mesh = meshes.icosphere(3, 0.03, 5)
    o = Oracle("oracle")
    center, radius = o.bounds(mesh)
    gate = EvalSettings(profile=0, screen_size=32, limit=1.25, supersample=2, max_supersample=8, orthographic=4,
                        perspective=2, rotation_seed=5)
    with Audit(0) as audit, Scorer(he_weights(4)) as scorer:
        def scores(indices, legal):
            g = scorer.graph(Mesh(mesh.positions, indices), center, radius)
            logits = scorer.forward(g, BF16)[:, 0]
            key = {(int(a), int(b)): i for i, (a, b) in enumerate(g.actions_array())}
            g.release()
            return np.array([logits[key[(int(u), int(v))]] for u, v in legal], np.float32)
        for batch in (4, 64):
            want = execute_actions(mesh, mesh.triangles // 3, 40, 0, batch, oracle_gate(mesh, center, radius, gate),
                                   scores=scores)
            out, stats = gpu_execute(audit, mesh, center, radius, gate, mesh.triangles // 3, 40, 0, batch,
                                     scorer=scorer)
            got = {k: stats[k] for k in ("ranked", "trials", "accepted", "rejected")}
            assert np.array_equal(out, want[0]) and got == want[1], (batch, got, want[1])
