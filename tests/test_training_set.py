"""Scorer training across many meshes: the uneven mesh set, teacher labels
(training/runtime_teacher.hpp:125-199, training/teacher_labels.hpp:16-36)
and joined device graphs.

CPU: preferred_actions against hand-worked pools (fewest faces before
smallest margin, exact ties, unconfirmed and non-finite choices skipped, the
face floor, the 16-row cap); the row label bits and margin of hand-built
measurements; the teacher row selection; the level gates against the
chain's schedule; the mesh set's sizes, families and determinism, and the
latitude-longitude sphere's closed-surface invariants.
GPU: a joined graph equals its parts (actions against the union's own
topology, features and FP32 / BF16 outputs bit for bit); labels of a small
mesh subset are consistent with their audits; and a training step on a batch
spanning meshes matches the FP64 backward of the union graph.
"""
from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "oracle"))

import scorer_oracle as so  # noqa: E402
import train_oracle as to  # noqa: E402
from blitzremesher_amd import meshes  # noqa: E402
from blitzremesher_amd.audit import Measurement  # noqa: E402
from blitzremesher_amd.teacher import (ADJACENT_KNOWN, ADJACENT_PASS, POOL, SOURCE_KNOWN, SOURCE_PASS,  # noqa: E402
                                       Choice, pool_rows, preferred_actions, row_label)


# --------------------------------------------------------------- labels (CPU)

def test_preferred_actions_fewest_faces_then_margin():
    c = [Choice(10, 0.1, True), Choice(9, 0.7, True), Choice(9, 0.5, True), Choice(9, 0.5, True)]
    assert preferred_actions(c) == 0b1100  # 9 faces beat 10; 0.5 beats 0.7; exact ties both
    # Unconfirmed, non-finite and below-floor choices never count.
    c = [Choice(5, 0.0, False), Choice(6, math.inf, True), Choice(6, math.nan, True), Choice(7, 0.9, True)]
    assert preferred_actions(c) == 0b1000
    c = [Choice(8, 0.2, True), Choice(12, 0.3, True)]
    assert preferred_actions(c, minimum_faces=10) == 0b10
    assert preferred_actions([Choice()] * 4) == 0
    with pytest.raises(ValueError):
        preferred_actions([Choice(1, 0, True)] * 17)


def measurement(error=0.0, area=0.0, passed=True, complete=True, views=16) -> Measurement:
    return Measurement(error=error, changed_area=area, passed=passed, complete=complete, views_evaluated=views)


def gates(limit=1.0, area=1.0):
    from blitzremesher_amd.audit import EvalSettings
    return tuple(EvalSettings(limit=limit, max_changed_area=area) for _ in range(4))


def test_row_label_bits_and_margin():
    g = gates(2.0, 0.5)
    ok = measurement(0.5, 0.1)
    label, choice = row_label(ok, measurement(1.0, 0.2), ok, ok, g, 100)
    assert label == SOURCE_KNOWN | ADJACENT_KNOWN | SOURCE_PASS | ADJACENT_PASS
    assert choice.confirmed and choice.faces == 100 and choice.margin == 0.5  # max(1.0 / 2, 0.2 / 0.5 = 0.4)
    # A failure witnessed over its limit is known; the source does not pass.
    label, choice = row_label(measurement(3.0, passed=False), ok, ok, ok, g, 100)
    assert label == SOURCE_KNOWN | ADJACENT_KNOWN | ADJACENT_PASS and not choice.confirmed
    # A failure under both limits, or a cancelled audit, is unknown.
    label, _ = row_label(ok, measurement(1.0, 0.1, passed=False), ok, ok, g, 100)
    assert label == SOURCE_KNOWN | SOURCE_PASS
    label, _ = row_label(ok, ok, Measurement(cancelled=True), ok, g, 100)
    assert label == ADJACENT_KNOWN | ADJACENT_PASS
    # An incomplete pass is known to nobody; a nan error is never known.
    label, _ = row_label(measurement(complete=False), ok, ok, ok, g, 100)
    assert not label & SOURCE_KNOWN
    label, _ = row_label(measurement(math.nan, passed=False), ok, ok, ok, g, 100)
    assert not label & SOURCE_KNOWN
    # Four equal passes: the margin is the larger of 0.5 / 2 and 0.1 / 0.5.
    _, choice = row_label(ok, ok, ok, ok, g, 90)
    assert choice.confirmed and choice.margin == 0.25
    # A zero area limit: zero changed area has ratio 0, any other infinity.
    _, choice = row_label(ok, ok, ok, ok, gates(2.0, 0.0), 90)
    assert choice.margin == math.inf


def test_pool_rows_take_the_classical_head_then_random():
    order = np.random.default_rng(1).permutation(500).astype(np.uint32)
    rows = pool_rows(order, 20, 3)
    assert rows.shape == (20, POOL)
    for p in range(20):
        assert np.array_equal(rows[p, :8], order[8 * p:8 * p + 8])
        assert len(set(rows[p].tolist())) == POOL
    assert np.array_equal(rows, pool_rows(order, 20, 3))
    assert not np.array_equal(rows, pool_rows(order, 20, 4))
    with pytest.raises(ValueError):
        pool_rows(order[:15], 1, 0)


def test_level_gates_follow_the_schedule():
    from blitzremesher_amd.chain import ChainSettings, schedule
    from blitzremesher_amd.teacher import level_gates
    s = ChainSettings(levels=4, profile=0, base_pixels=256, last_pixels=32, max_lod0_delta_px=1.0,
                      transition=((0, 1.5), (1, 3)), search_orthographic=16, search_perspective=2,
                      audit_orthographic=24, audit_perspective=0, search_supersample=4, audit_supersample=8,
                      max_supersample=16, max_changed_area=0.25, normal_importance=((0, 1), (1, 3)))
    center, radius = np.zeros(3), 1.0
    pixels, transition, source = schedule(s, center, radius)[2]
    sa, sb, a, b = level_gates(s, center, radius, level=2)
    assert (sa.limit, sb.limit, a.limit, b.limit) == (source, transition, source, transition)
    assert sa.screen_size == a.screen_size == pixels
    assert (sa.orthographic, sa.perspective, sa.rotation_seed, sa.supersample) == (16, 2, s.search_seed, 4)
    assert (a.orthographic, a.perspective, a.rotation_seed, a.supersample) == (24, 0, s.audit_seed, 8)
    assert sa.max_changed_area == 1.0 and a.max_changed_area == 0.25
    assert a.normal_weight == pytest.approx(s.normal_weight * 2.0)  # t = 1/2 on the curve 1 -> 3


# ------------------------------------------------------------ meshes (CPU)

def closed_surface(mesh) -> tuple[bool, int]:
    f = mesh.indices.astype(np.int64)
    e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), 1)
    _, counts = np.unique(e, axis=0, return_counts=True)
    return bool((counts == 2).all()), len(mesh.positions) - len(counts) + len(f)


def test_uv_sphere_is_a_closed_sphere():
    m = meshes.uv_sphere(7, 12, 0.0)
    closed, euler = closed_surface(m)
    assert closed and euler == 2
    assert m.triangles == 2 * 12 * 6 and len(m.positions) == 12 * 6 + 2
    assert np.allclose(np.linalg.norm(m.positions, axis=1), 1.0, atol=1e-6)
    # Outward winding: every face normal points away from the centre.
    p = m.positions.astype(np.float64)
    f = m.indices
    n = np.cross(p[f[:, 1]] - p[f[:, 0]], p[f[:, 2]] - p[f[:, 0]])
    assert ((n * (p[f[:, 0]] + p[f[:, 1]] + p[f[:, 2]])).sum(1) > 0).all()


def test_training_set_sizes_and_families():
    a = meshes.training_set(16, 2000, 20000, 5)
    b = meshes.training_set(16, 2000, 20000, 5)
    assert [f for f, _ in a] == list(meshes.TRAINING_FAMILIES) * 4
    for (fa, ma), (_, mb) in zip(a, b):
        assert np.array_equal(ma.positions, mb.positions) and np.array_equal(ma.indices, mb.indices)
    u = meshes.uniform(5, 40, 16)
    want = np.exp(np.log(2000) + u * (np.log(20000) - np.log(2000)))
    got = np.array([m.triangles for _, m in a])
    # Each family's lattice rounds its size: within 8% of the drawn count.
    assert (np.abs(got / want - 1) < 0.08).all(), np.round(got / want, 3)
    for family, m in a:
        closed, euler = closed_surface(m)
        assert closed == (family != "heightfield")
        if family == "sphere":
            assert euler == 2
        if family == "torus":
            assert euler == 0


# ----------------------------------------------------------------------- GPU

SUBSET = [("sphere", 900, 51), ("torus", 1400, 52), ("heightfield", 700, 53), ("extruded_box", 1100, 54)]


def subset():
    return [meshes.training_mesh(f, n, s) for f, n, s in SUBSET]


def union(ms):
    idx, base = [], 0
    for m in ms:
        idx.append(m.indices.astype(np.int64) + base)
        base += len(m.positions)
    return np.concatenate(idx), base


@pytest.mark.gpu
def test_gpu_joined_graph_equals_its_parts():
    from blitzremesher_amd.audit import bounds
    from blitzremesher_amd.scorer import BF16, FP32, Scorer, he_weights
    ms = subset()
    with Scorer(he_weights(4)) as scorer:
        parts = []
        for m in ms:
            c, r = bounds(m)
            parts.append(scorer.graph(m, c, r))
        joined = scorer.concat(parts)
        idx, nv = union(ms)
        offsets, neighbors, *_ = so.topology(idx, nv)
        assert joined.vertices == nv and joined.actions == len(neighbors)
        assert np.array_equal(joined.actions_array(), so.actions(offsets, neighbors))
        feats = np.concatenate([g.features() for g in parts])
        assert np.array_equal(joined.features().view(np.uint32), feats.view(np.uint32))
        outputs = {}
        for precision in (FP32, BF16):
            want = np.concatenate([scorer.forward(g, precision) for g in parts])
            outputs[precision] = scorer.forward(joined, precision)
            assert np.array_equal(outputs[precision].view(np.uint32), want.view(np.uint32)), precision
        for g in parts:
            g.release()
        # The joined graph owns its buffers: it outlives its parts.
        again = scorer.forward(joined, FP32)
        assert np.array_equal(again.view(np.uint32), outputs[FP32].view(np.uint32))


@pytest.fixture(scope="module")
def subset_dataset():
    from bench import c1_chain_settings
    from blitzremesher_amd import Audit
    from blitzremesher_amd.scorer import Scorer, he_weights
    from blitzremesher_amd.teacher import build_dataset
    ms = subset()
    audit = Audit(0)
    scorer = Scorer(he_weights(4))
    data = build_dataset(audit, scorer, ms, c1_chain_settings(), 6, seed=5)
    yield ms, audit, scorer, data
    scorer.close()
    audit.close()


@pytest.mark.gpu
def test_gpu_subset_labels_follow_their_audits(subset_dataset):
    from blitzremesher_amd.scorer import PLACED, PREFERRED, VALID
    from blitzremesher_amd.teacher import PLACEMENT_PREFERRED
    ms, audit, scorer, data = subset_dataset
    assert data.pools == 6 * len(ms) and len(data.actions) == data.pools * POOL
    assert np.array_equal(data.pool_mesh, np.repeat(np.arange(len(ms)), 6))
    acts = data.graph.actions_array()
    base = 0
    for k, (m, lab) in enumerate(zip(ms, data.meshes)):
        rows = slice(k * 6 * POOL, (k + 1) * 6 * POOL)
        # Joined action ids name the labelled collapses, offset per mesh.
        assert np.array_equal(acts[data.actions[rows]] - base, lab.pairs)
        base += len(m.positions)
        for p in range(6):
            r = slice(p * POOL, (p + 1) * POOL)
            pref = (lab.label[r] & PLACEMENT_PREFERRED) != 0
            conf = np.isfinite(lab.margin[r])
            # Preferred rows are confirmed, share the fewest faces and then the
            # smallest margin among confirmed rows.
            assert not (pref & ~conf).any()
            if conf.any():
                faces = lab.faces[r][conf].min()
                best = lab.margin[r][conf & (lab.faces[r] == faces)].min()
                assert np.array_equal(pref, conf & (lab.faces[r] == faces) & (lab.margin[r] == best))
    bits = data.bits
    assert ((bits & VALID) != 0).sum() > data.pools * 8
    assert ((bits & PREFERRED) != 0).any() and ((bits & PLACED) != 0).any()
    assert not (data.targets[(bits & PLACED) == 0]).any()
    assert (np.abs(data.targets) <= 1).all()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_gpu_batch_spanning_meshes_matches_fp64(subset_dataset, precision):
    from blitzremesher_amd.scorer import TrainSettings, he_weights
    ms, audit, scorer, data = subset_dataset
    w = he_weights(4)
    scorer.set_weights(w)
    scorer.reset_optimizer()
    rng = np.random.default_rng(3)
    take = rng.permutation(data.pools)[:16]
    rows = (take[:, None] * POOL + np.arange(POOL)).reshape(-1)
    assert len(set(data.pool_mesh[take].tolist())) == len(ms)  # the batch spans every mesh
    ids, bits, targets = data.actions[rows], data.bits[rows], data.targets[rows]
    r = scorer.train_step(data.graph, ids, bits, targets, TrainSettings(pool=POOL, precision=precision))
    grad = scorer.gradients()
    idx, nv = union(ms)
    offsets, neighbors, *_ = so.topology(idx, nv)
    acts = so.actions(offsets, neighbors)
    loss, want = to.backward_fp64(data.graph.features(), offsets, neighbors, acts[ids], w.astype(np.float64), bits,
                                  targets, POOL)
    errs = []
    for mat, bias in to.layer_slices():
        for sl in (mat, bias):
            errs.append(np.linalg.norm(grad[sl] - want[sl]) / max(np.linalg.norm(want[sl]), 1e-12))
    tol = 1e-5 if precision == 0 else 0.1
    print(f"precision {precision}: loss {r.loss:.6f} vs {loss:.6f}; layer errors {np.round(errs, 8)}")
    assert r.applied
    assert abs(r.loss - loss) <= tol * max(1.0, abs(loss))
    assert max(errs) <= tol
