"""Edge-graph scorer: oracle pins (CPU) and HIP parity (GPU).

The FP64 oracle (oracle/scorer_oracle.py) is pinned by closed forms that
follow the reference's own scorer tests (tests/neural.cpp:82-117: bias-only
output, directed neighbour mean, isolated vertex) and by an independent
per-vertex loop for the features. GPU checks: features and the FP32 mode are
bit-exact against the oracle's restatements; the BF16 mode is within a stated
tolerance of FP64; rankings and placements match exactly.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "oracle"))

import scorer_oracle as so  # noqa: E402
from oracle_lib import available  # noqa: E402
from blitzremesher_amd import meshes  # noqa: E402
from blitzremesher_amd.audit import Mesh, bounds  # noqa: E402


needs_ref = pytest.mark.skipif(not available("ref"), reason="reference CPU build (oracle/_ref) not present")


def tiny_graph():
    # Directed-mean semantics of tests/neural.cpp:99-113 on an undirected
    # graph: path 0-1-2 plus isolated vertex 3.
    offsets = np.array([0, 1, 3, 4, 4])
    neighbors = np.array([1, 0, 2, 1])
    return offsets, neighbors


def test_weight_layout():
    from blitzremesher_amd.scorer import he_weights, weight_count
    w = he_weights(4)
    assert w.size == so.weight_count() == weight_count() == 107524
    for (W, b), (o, i) in zip(so.split_weights(w), so.SHAPES):
        assert W.shape == (o, i) and not b.any()
        assert abs(W.std() - np.sqrt(2.0 / i)) < 0.1 * np.sqrt(2.0 / i)


def test_oracle_bias_only_closed_form():
    # All matrices zero: every output equals the last layer's bias.
    w = np.zeros(so.weight_count(), np.float32)
    layers = so.split_weights(w)
    for k, (W, b) in enumerate(layers):
        b[:] = k + 1
    offsets, neighbors = tiny_graph()
    x = np.random.default_rng(0).random((4, so.FEATURES))
    acts = so.actions(offsets, neighbors)
    out = so.forward_fp64(x, offsets, neighbors, acts, w)
    assert np.all(out == 5.0)
    assert np.array_equal(so.forward_fp32(x, offsets, neighbors, acts, w), out.astype(np.float32))


def test_oracle_neighbour_mean_closed_form():
    # Layer 0 copies the mean of feature 0 into channel 0, later layers pass
    # channel 0 through, the edge layer copies h(dst) channel 0, head = it.
    w = np.zeros(so.weight_count(), np.float32)
    layers = so.split_weights(w)
    layers[0][0][0, so.FEATURES] = 1.0
    layers[1][0][0, 0] = 1.0
    layers[2][0][0, 0] = 1.0
    layers[3][0][0, so.HIDDEN] = 1.0
    layers[4][0][0, 0] = 1.0
    offsets, neighbors = tiny_graph()
    x = np.zeros((4, so.FEATURES))
    x[:, 0] = [2, 4, 8, 16]
    acts = so.actions(offsets, neighbors)
    out = so.forward_fp64(x, offsets, neighbors, acts, w)
    # h1 channel 0 = neighbour means: v0 -> 4, v1 -> (2+8)/2 = 5, v2 -> 4, v3 -> 0 (isolated).
    h = {0: 4.0, 1: 5.0, 2: 4.0, 3: 0.0}
    assert [out[a, 0] for a in range(len(acts))] == [h[int(d)] for d in acts[:, 1]]
    assert np.array_equal(so.forward_fp32(x, offsets, neighbors, acts, w)[:, 0], out[:, 0].astype(np.float32))


def test_compensated_sum_recovers_cancellation():
    # Compensated FP32 keeps terms a naive float32 sum loses.
    # Lane 0 sees 1e8 followed by eight ones (float spacing at 1e8 is 8).
    w = np.zeros((1, 72), np.float32)
    x = np.zeros((1, 72), np.float32)
    x[0, 0::8] = [1e8] + [1.0] * 8
    w[0, 0::8] = 1.0
    got = so._compensated_linear(x, w, np.zeros(1, np.float32), False)
    assert got[0, 0] == np.float32(100000008.0)
    naive = np.float32(1e8)
    for _ in range(8):
        naive = np.float32(naive + np.float32(1.0))
    assert naive == np.float32(1e8)


def test_oracle_fp32_close_to_fp64():
    from blitzremesher_amd.scorer import he_weights
    m = meshes.torus(1.0, 0.3, 12, 8, 0.01, 3)
    c, r = bounds(m)
    x = so.features(m.positions, m.indices, c, 2 * r).astype(np.float32)
    offsets, neighbors, *_ = so.topology(m.indices, len(m.positions))
    acts = so.actions(offsets, neighbors)
    w = he_weights(4)
    ref = so.forward_fp64(x, offsets, neighbors, acts, w)
    f32 = so.forward_fp32(x, offsets, neighbors, acts, w)
    assert np.abs(f32 - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())


def test_features_closed_form_on_flat_grid():
    m = meshes.grid(4, double_sided=False)
    c, r = bounds(m)
    x = so.features(m.positions, m.indices, c, 2 * r)
    assert np.allclose(x[:, 3:6], [0, 0, 1])
    assert np.allclose(x[:, 16], 0)
    # Plane z = 0: only a22 (n_z^2) is nonzero after area normalisation.
    q = x[:, 6:16]
    assert np.allclose(q[:, 7], 1) and np.allclose(np.delete(q, 7, axis=1), 0)
    # Corner (0,0): neighbours (1,0), (0,1), (1,1); two boundary edges.
    assert x[0, 17] == 3 / 16 and x[0, 22] == 2 / 4 and x[0, 18] == 1
    centre = 2 * 5 + 2
    assert x[centre, 17] == 6 / 16 and x[centre, 18] == 0


def bipyramid(k: int, r: float, h: float) -> Mesh:
    """A regular k-gon of radius r in z = 0 (vertices 0..k-1, counter-clockwise
    from above) joined to apexes (0, 0, h) (vertex k) and (0, 0, -h) (vertex
    k+1); closed, faces wound outward."""
    t = 2 * np.pi * np.arange(k) / k
    p = np.zeros((k + 2, 3))
    p[:k, 0], p[:k, 1] = r * np.cos(t), r * np.sin(t)
    p[k, 2], p[k + 1, 2] = h, -h
    f = [(k, i, (i + 1) % k) for i in range(k)] + [(k + 1, (i + 1) % k, i) for i in range(k)]
    return Mesh(p.astype(np.float32), np.array(f, np.uint32))


@pytest.mark.parametrize("k,r,h", [(6, 1.0, 0.5), (8, 1.0, 0.25), (10, 0.75, 1.5), (12, 2.0, 0.1)])
def test_features_closed_form_on_curved_apexes(k, r, h):
    # Each apex sits h above (below) a regular ring of radius r: its
    # neighbour mean is the ring centre, so lap = (0, 0, -+h) / diameter, the
    # area-weighted normal is (0, 0, +-1), every incident edge has length
    # sqrt(r^2 + h^2), and every incident face is tilted by alpha with
    # tan(alpha) = h / (r cos(pi/k)) and has |cross| = 2 r sin(pi/k)
    # sqrt(r^2 cos^2(pi/k) + h^2). The bounds are centred at the origin with
    # radius max(r, h).
    m = bipyramid(k, r, h)
    c, radius = bounds(m)
    assert np.allclose(c, 0, atol=1e-6) and abs(radius - max(r, h)) < 1e-6 * max(r, h)
    diameter = 2 * radius
    x = so.features(m.positions, m.indices, c, diameter)
    slant = np.sqrt(r * r + h * h)
    mean_edge = slant / diameter
    alpha = np.arctan2(h, r * np.cos(np.pi / k))
    total = k * 2 * r * np.sin(np.pi / k) * np.sqrt((r * np.cos(np.pi / k)) ** 2 + h * h) / diameter ** 2
    for apex, up in ((k, 1.0), (k + 1, -1.0)):
        a = x[apex]
        assert np.allclose(a[3:6], [0, 0, up], atol=1e-6)
        assert a[16] == pytest.approx(-h / slant, rel=1e-6)                     # curvature
        assert a[17] == k / 16 and a[21] == k / 8 and a[18] == 0 and a[22] == 0  # valence, faces, boundary
        assert a[20] == pytest.approx(min(64 * mean_edge, 4.0), rel=1e-6)         # mean edge
        assert a[30] == pytest.approx(min(np.sqrt(total) / mean_edge, 4.0), rel=1e-6)
        q = a[6:16]  # xx xy xz xw yy yz yw zz zw ww
        d = -np.cos(alpha) * h / diameter  # plane offset of every incident face
        want = [np.sin(alpha) ** 2 / 2, 0, 0, 0, np.sin(alpha) ** 2 / 2, 0, 0, np.cos(alpha) ** 2,
                up * np.cos(alpha) * d, d * d]
        assert np.allclose(q, want, atol=1e-6), (q, want)
    ring = x[:k]
    assert np.all(ring[:, 17] == 4 / 16) and np.all(ring[:, 21] == 4 / 8)
    assert np.allclose(ring[:, 3:6] @ np.array([0, 0, 1.0]), 0, atol=1e-6)  # ring normals horizontal


@needs_ref
def test_oracle_scores_match_reference_build():
    # The shortest-edge and current-plane orders rank by the reference's own
    # scores: x[59] of its action rows and ActionState::teacher_cost.
    from oracle_lib import Actions, Oracle
    o = Oracle("ref")
    for mesh in (meshes.icosphere(2, 0.02, 1), meshes.torus(1.0, 0.3, 12, 8, 0.02, 3),
                 meshes.jitter(meshes.heightfield(10, 0.3, 5), 0.01, 5, 0.5), bipyramid(7, 1.0, 0.6)):
        _, radius = o.bounds(mesh)
        acts = Actions("ref", mesh)
        legal = acts.legal.astype(np.int64)
        assert len(legal) > 0
        for mode in (3, 4):
            want = acts.ref_scores(mode) + np.float32(0)
            got = so.scores(mesh.positions, mesh.indices, legal, mode, diameter=max(1e-20, 2 * radius))
            assert np.array_equal(got, want), (mode, np.abs(got - want).max())
        acts.close()


def test_topology_is_symmetric_with_reverse():
    m = meshes.icosphere(2, 0.02, 1)
    offsets, neighbors, face_offsets, faces, boundary = so.topology(m.indices, len(m.positions))
    acts = so.actions(offsets, neighbors)
    pairs = {(int(a), int(b)) for a, b in acts}
    assert all((b, a) in pairs for a, b in pairs)
    assert len(boundary) == 0 and np.all(np.diff(offsets) >= 5)
    assert np.array_equal(np.diff(face_offsets), np.bincount(m.indices.reshape(-1), minlength=len(m.positions)))


# ----------------------------------------------------------------------- GPU

@pytest.fixture(scope="module")
def scorer():
    from blitzremesher_amd.scorer import Scorer, he_weights
    s = Scorer(he_weights(4))
    yield s
    s.close()


def graph_case(scorer, mesh):
    c, r = bounds(mesh)
    g = scorer.graph(mesh, c, r)
    offsets, neighbors, *_ = so.topology(mesh.indices, len(mesh.positions))
    acts = so.actions(offsets, neighbors)
    return g, c, r, offsets, neighbors, acts


CASES = {
    "torus": lambda: meshes.torus(1.0, 0.3, 40, 16, 0.01, 3),
    "icosphere": lambda: meshes.icosphere(3, 0.02, 1),
    "heightfield": lambda: meshes.heightfield(24, 0.1, 2),
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_features_and_actions_match_oracle(scorer, case):
    mesh = CASES[case]()
    g, c, r, offsets, neighbors, acts = graph_case(scorer, mesh)
    assert np.array_equal(g.actions_array(), acts)
    want = so.features(mesh.positions, mesh.indices, c, 2 * r).astype(np.float32)
    assert np.array_equal(g.features(), want)


def fan(k: int) -> Mesh:
    """An open disc: centre vertex 0 of valence k joined to a rim of k
    vertices (every rim edge a boundary edge)."""
    t = 2 * np.pi * np.arange(k) / k
    p = np.zeros((k + 1, 3), np.float32)
    p[1:, 0], p[1:, 1], p[1:, 2] = np.cos(t), np.sin(t), 0.05 * np.sin(3 * t)
    f = [(0, 1 + i, 2 + i) for i in range(k - 1)]
    return Mesh(p, np.array(f, np.uint32))


def ragged() -> Mesh:
    """A heightfield with unused vertices in its vertex stream (before,
    between and after the used ones), faces with a repeated corner and a
    face of one vertex repeated."""
    base = meshes.heightfield(10, 0.1, 4)
    nv = len(base.positions)
    rng = np.random.default_rng(9)
    extra = rng.normal(0, 1, (7, 3)).astype(np.float32)
    positions = np.concatenate([extra[:2], base.positions[: nv // 2], extra[2:4], base.positions[nv // 2:], extra[4:]])
    remap = np.concatenate([np.arange(nv // 2) + 2, np.arange(nv // 2, nv) + 4]).astype(np.uint32)
    faces = remap[base.indices]
    degenerate = np.array([[faces[3, 0], faces[3, 0], faces[3, 1]], [faces[7, 2], faces[9, 1], faces[9, 1]],
                           [faces[11, 0], faces[11, 0], faces[11, 0]]], np.uint32)
    return Mesh(positions, np.concatenate([faces[:20], degenerate, faces[20:]]).astype(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["fan", "ragged", "one_face", "empty"])
def test_gpu_device_topology_matches_oracle(scorer, case):
    """The device-built topology (edge sort, run-length counts, directed
    sort, row starts) gives the oracle's actions on meshes with a vertex of
    valence 300, an open boundary, vertices no face uses (before, between
    and after used ones), degenerate faces, one face and no faces."""
    if case == "fan":
        mesh = fan(300)
    elif case == "ragged":
        mesh = ragged()
    elif case == "one_face":
        mesh = Mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [5, 5, 5]], np.float32), np.array([[0, 1, 2]], np.uint32))
    else:
        mesh = Mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.zeros((0, 3), np.uint32))
    g, c, r, offsets, neighbors, acts = graph_case(scorer, mesh)
    assert np.array_equal(g.actions_array(), acts.astype(np.uint32).reshape(-1, 2))
    if case == "fan":
        assert np.diff(offsets)[0] == 300
        want = so.features(mesh.positions, mesh.indices, c, 2 * r).astype(np.float32)
        assert np.array_equal(g.features(), want)
    # A second build into the same graph (buffers reused) gives the same.
    g2 = scorer.graph(mesh, c, r)
    assert np.array_equal(g2.actions_array(), g.actions_array())


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_fp32_forward_is_bit_exact(scorer, case):
    from blitzremesher_amd.scorer import FP32
    from blitzremesher_amd.scorer import he_weights
    mesh = CASES[case]()
    g, c, r, offsets, neighbors, acts = graph_case(scorer, mesh)
    got = scorer.forward(g, FP32)
    want = so.forward_fp32(g.features(), offsets, neighbors, acts, he_weights(4))
    assert np.array_equal(got, want)


# BF16 tolerance: activations and weights rounded to 8-bit mantissas
# (relative 2^-9 each) through four layers of width <= 256 with FP32
# accumulation; measured errors are 6.7e-3 to 8.7e-3 of the output scale.
# The bound below is 2e-2 of the largest FP64 output magnitude.
BF16_TOLERANCE = 2e-2


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_bf16_forward_within_tolerance_of_fp64(scorer, case):
    from blitzremesher_amd.scorer import BF16, FP32, he_weights
    mesh = CASES[case]()
    g, c, r, offsets, neighbors, acts = graph_case(scorer, mesh)
    ref = so.forward_fp64(g.features(), offsets, neighbors, acts, he_weights(4))
    bf = scorer.forward(g, BF16)
    f32 = scorer.forward(g, FP32)
    scale = max(1.0, np.abs(ref).max())
    err_bf = np.abs(bf - ref).max() / scale
    err_32 = np.abs(f32 - ref).max() / scale
    print(f"{case}: bf16 max rel {err_bf:.3e} rms {np.sqrt(((bf - ref) ** 2).mean()) / scale:.3e}; "
          f"fp32 max rel {err_32:.3e}")
    assert err_32 <= 1e-5
    assert err_bf <= BF16_TOLERANCE


@pytest.mark.gpu
def test_gpu_rankings_match_oracle(scorer):
    from blitzremesher_amd.scorer import CONSTANT, CURRENT_PLANE, FP32, LEARNED, SHORTEST_EDGE
    mesh = meshes.torus(1.0, 0.3, 30, 12, 0.01, 3)
    g, c, r, offsets, neighbors, acts = graph_case(scorer, mesh)
    out = scorer.forward(g, FP32)
    for mode in (LEARNED, CONSTANT, SHORTEST_EDGE, CURRENT_PLANE):
        want = so.rank(so.scores(mesh.positions, mesh.indices, acts, mode, out, 2 * r))
        assert np.array_equal(scorer.rank(g, mode), want), mode


@pytest.mark.gpu
def test_gpu_decode_matches_oracle(scorer):
    from blitzremesher_amd.scorer import BF16
    mesh = meshes.icosphere(2, 0.02, 1)
    g, c, r, offsets, neighbors, acts = graph_case(scorer, mesh)
    out = scorer.forward(g, BF16)
    pick = np.arange(0, len(acts), 7, dtype=np.uint32)
    got = scorer.decode(g, pick)
    want = so.decode_placements(mesh.positions, acts[pick], out[pick])
    assert np.array_equal(got, want)


def corner_mesh():
    # Three unit quads on the planes x = 0, y = 0, z = 0 meeting at the
    # origin (vertex 0), two triangles each.
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [0, 1, 1], [1, 0, 1]], np.float32)
    idx = np.array([[0, 2, 4], [0, 4, 1],      # z = 0
                    [0, 1, 6], [0, 6, 3],      # y = 0
                    [0, 3, 5], [0, 5, 2]],     # x = 0
                   np.uint32)
    return p, idx


def test_teacher_oracle_closed_forms():
    # Corner: the plane quadrics of the faces around the origin and around
    # (1, 0, 0) vanish only at the origin; flat grid: the quadric is rank one
    # (singular) and the candidate falls back to the middle of the edge.
    p, idx = corner_mesh()
    offsets, neighbors, face_offsets, faces, _ = so.topology(idx, len(p))
    acts = np.array([[0, 1], [1, 0]], np.uint32)
    pos, tgt = so.teacher_placements(p, idx, face_offsets, faces, acts)
    assert np.allclose(pos[:, 3], 0, atol=1e-6)
    assert np.array_equal(pos[0, 0], p[0]) and np.array_equal(pos[0, 1], p[1])
    assert np.array_equal(pos[0, 2], [0.5, 0, 0])
    assert np.allclose(pos[0, 4:], [[0.75, 0, 0], [0.25, 0, 0], [0.5, 0.25, 0], [0.5, -0.25, 0], [0.5, 0, 0.25],
                                    [0.5, 0, -0.25]])
    assert np.allclose(tgt[0, 0], [-0.5, 0, 0]) and np.allclose(tgt[0, 4], [0.25, 0, 0])
    grid = meshes.grid(4)
    offsets, neighbors, face_offsets, faces, _ = so.topology(grid.indices, len(grid.positions))
    acts = so.actions(offsets, neighbors)
    pos, _ = so.teacher_placements(grid.positions, grid.indices, face_offsets, faces, acts[:40])
    assert np.array_equal(pos[:, 3], pos[:, 2])


def test_decode_oracle_closed_form():
    p = np.array([[0, 0, 0], [2, 0, 0]], np.float32)
    out = np.array([[9, 0.5, -3, 0.25]], np.float32)
    got = so.decode_placements(p, np.array([[0, 1]]), out)
    assert np.array_equal(got, np.array([[2.0, -2.0, 0.5]], np.float32))


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_teacher_matches_oracle(scorer, case):
    mesh = CASES[case]()
    g, c, r, offsets, neighbors, acts = graph_case(scorer, mesh)
    _, _, face_offsets, faces, _ = so.topology(mesh.indices, len(mesh.positions))
    pick = np.arange(0, len(acts), max(1, len(acts) // 300), dtype=np.uint32)
    got_p, got_t = scorer.teacher(g, pick)
    want_p, want_t = so.teacher_placements(mesh.positions, mesh.indices, face_offsets, faces, acts[pick])
    assert np.array_equal(got_p, want_p)
    assert np.array_equal(got_t, want_t)


@pytest.mark.gpu
def test_gpu_shuffled_ranking_is_unavailable(scorer):
    from blitzremesher_amd import BlitzError
    from blitzremesher_amd.scorer import SHUFFLED
    g = graph_case(scorer, meshes.icosphere(1, 0.0))[0]
    with pytest.raises(BlitzError) as e:
        scorer.rank(g, SHUFFLED)
    assert e.value.status == 6
