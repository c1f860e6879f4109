"""Action and placement policies (ActionCuda, src/neural/action.hpp:70-87,
src/neural/inference.cu:183-290), the learned action order built on them
(src/neural/action_gpu.cu:1751-1798) and placement decoding with grid
snapping (src/neural/action_gpu.cu:1157-1200, 374-390).

CPU: the oracle's policy restatement (three compensated layers) reproduces a
selection policy whose first output is minus the action's normalised edge
length (channel 59) exactly, and catches a wrong layer layout; the unorm16
grid restatement meets its closed forms (ends, half-up rounding, clamping,
zero extent, idempotence, outside points kept).
GPU: predictions equal the restatement bit for bit for each architecture
and hidden width, on real feature rows and random rows; the learned order of
the selection policy is the shortest-edge order, and random policies order
as the restatement's outputs; bad weights and inputs are rejected;
placements (supervised normals or not, snapped or not, exact vertices) equal
the restatement, default placements too, and a zero policy places at the
edge midpoint.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent))
sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
import scorer_oracle as so  # noqa: E402
from feature_cases import CONDITIONS  # noqa: E402

from blitzremesher_amd.actions import POLICY_SHAPES, policy_weight_count  # noqa: E402

FIXTURE = Path(__file__).resolve().parent / "golden" / "action_features.npz"


def random_weights(architecture: int, hidden: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    i, o = POLICY_SHAPES[architecture]
    parts = []
    for k_in, k_out in ((i, hidden), (hidden, hidden), (hidden, o)):
        parts.append((rng.standard_normal((k_out, k_in)) * np.sqrt(2.0 / k_in)).reshape(-1))
        parts.append(rng.standard_normal(k_out) * 0.1)
    w = np.concatenate(parts).astype(np.float32)
    assert w.size == policy_weight_count(architecture, hidden)
    return w


def selection_weights(hidden: int = 64) -> np.ndarray:
    """Hidden unit 0 = relu(x[59]) (the normalised edge length, >= 0), passed
    through; output 0 = -h0, outputs 1, 2 = 0."""
    i, o = POLICY_SHAPES[2]
    w0 = np.zeros((hidden, i), np.float32)
    w0[0, 59] = 1
    w1 = np.zeros((hidden, hidden), np.float32)
    w1[0, 0] = 1
    w2 = np.zeros((o, hidden), np.float32)
    w2[0, 0] = -1
    return np.concatenate([w0.ravel(), np.zeros(hidden, np.float32), w1.ravel(), np.zeros(hidden, np.float32),
                           w2.ravel(), np.zeros(o, np.float32)])


def fixture_rows(width: int = 80) -> np.ndarray:
    data = np.load(FIXTURE)
    rows = np.concatenate([data[k] for k in data.files if k.endswith("/rows")])
    if width == 128:
        rows = np.concatenate([rows, np.zeros((rows.shape[0], 48), np.float32)], 1)
    return rows


def test_oracle_selection_policy_is_minus_edge_length():
    rows = fixture_rows()
    out = so.policy_forward(rows, selection_weights(), 64, 3)
    assert np.array_equal(out[:, 0], -rows[:, 59])
    assert not out[:, 1:].any()


def test_oracle_policy_catches_a_transposed_layer():
    # The first layer read as [in, out] instead of [out, in] changes outputs.
    rows = fixture_rows()[:64]
    w = random_weights(2, 64, 3)
    want = so.policy_forward(rows, w, 64, 3)
    bad = w.copy()
    bad[:80 * 64] = w[:80 * 64].reshape(64, 80).T.reshape(-1)
    assert not np.array_equal(so.policy_forward(rows, bad, 64, 3), want)


# ----------------------------------------------------------------------- GPU

@pytest.mark.gpu
@pytest.mark.parametrize("architecture,hidden", [(2, 64), (2, 256), (3, 128), (4, 64)])
def test_gpu_policy_matches_restatement(architecture, hidden):
    from blitzremesher_amd.actions import Policy
    width, outputs = POLICY_SHAPES[architecture]
    rng = np.random.default_rng(architecture * 1000 + hidden)
    rows = np.concatenate([fixture_rows(width), rng.standard_normal((300, width)).astype(np.float32)])
    w = random_weights(architecture, hidden, hidden + architecture)
    with Policy(w, architecture, hidden) as p:
        got = p.predict(rows)
    want = so.policy_forward(rows, w, hidden, outputs)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), np.abs(got - want).max()


@pytest.mark.gpu
def test_gpu_selection_policy_ranks_as_shortest_edge():
    from blitzremesher_amd import meshes
    from blitzremesher_amd.actions import ActionSet, Policy
    mesh = meshes.with_attributes(meshes.icosphere(3, 0.02, 4), 2)
    with ActionSet(mesh) as s, Policy(selection_weights()) as p:
        learned = s.rank_policy(p, CONDITIONS)
        shortest = s.rank(3)
    assert np.array_equal(learned, shortest)


@pytest.mark.gpu
@pytest.mark.parametrize("architecture", [2, 3])
def test_gpu_policy_order_matches_restatement(architecture):
    from blitzremesher_amd import meshes
    from blitzremesher_amd.actions import ActionSet, Policy
    width, outputs = POLICY_SHAPES[architecture]
    mesh = meshes.with_attributes(meshes.icosphere(3, 0.02, 6), 3)
    w = random_weights(architecture, 64, 17)
    with ActionSet(mesh) as s, Policy(w, architecture, 64) as p:
        rows = s.features(CONDITIONS, width=width)
        order = s.rank_policy(p, CONDITIONS)
    score = so.policy_forward(rows, w, 64, outputs)[:, 0] + np.float32(0)
    want = np.lexsort((np.arange(len(score)), -score.astype(np.float64)))
    assert np.array_equal(order, want)


@pytest.mark.gpu
def test_gpu_policy_rejects_bad_weights_and_inputs():
    from blitzremesher_amd._lib import BlitzError
    from blitzremesher_amd.actions import Policy
    w = random_weights(2, 64, 1)
    with pytest.raises(BlitzError):
        Policy(w[:-1], 2, 64)
    with pytest.raises(BlitzError):
        Policy(w, 2, 96)
    bad = w.copy()
    bad[5] = np.nan
    with pytest.raises(BlitzError):
        Policy(bad, 2, 64)
    with Policy(w, 2, 64) as p:
        x = np.zeros((2, 80), np.float32)
        x[1, 3] = np.inf
        with pytest.raises(BlitzError):
            p.predict(x)


# ------------------------------------------------------------- placements

def test_oracle_unorm16_grid_closed_forms():
    # Box [0, 2]: grid step 2/65535. Ends, a midpoint between two steps
    # (rounded half up), clamping and zero extent.
    assert so.pack_unorm16(0.0, 0.0, 2.0) == 0 and so.pack_unorm16(2.0, 0.0, 2.0) == 65535
    assert so.pack_unorm16(-1.0, 0.0, 2.0) == 0 and so.pack_unorm16(3.0, 0.0, 2.0) == 65535
    assert so.pack_unorm16(1.0, 0.0, 2.0) == 32768  # 32767.5 rounds up
    assert so.pack_unorm16(5.0, 1.0, 0.0) == 0
    assert so.unpack_unorm16(65535, 0.0, 2.0) == np.float32(2.0)
    assert so.unpack_unorm16(0, -1.5, 2.0) == np.float32(-1.5)
    # Snapping is idempotent and outside points are kept.
    q = ((-1, -1, -1), (2, 2, 2))
    p = np.array([0.123456, -0.5, 0.75], np.float32)
    g = so.grid_position(p, *q)
    assert np.array_equal(so.grid_position(g, *q), g) and not np.array_equal(g, p)
    out = np.array([1.5, 0, 0], np.float32)
    assert np.array_equal(so.grid_position(out, *q), out)


def wedge_checks(mesh, legal, mapping):
    pos = mesh.positions
    for (u, v), m in zip(legal, mapping):
        assert m[0] in (1, 2)
        for j in range(m[0]):
            s, t = m[1 + 2 * j], m[2 + 2 * j]
            assert np.array_equal(pos[s], pos[u]) and np.array_equal(pos[t], pos[v])
        if m[0] == 2:
            assert (m[1], m[2]) < (m[3], m[4])


@pytest.mark.gpu
@pytest.mark.parametrize("supervise,quantized,exact", [(False, False, False), (True, False, False),
                                                       (True, True, False), (True, True, True)])
def test_gpu_placements_match_restatement(supervise, quantized, exact):
    from blitzremesher_amd import meshes
    from blitzremesher_amd.actions import ActionSet, Policy
    from feature_cases import seamed
    mesh, uv, tangents = seamed(meshes.icosphere(3, 0.02, 6), 5)
    c = CONDITIONS.copy()
    if not supervise:
        c[3:6] = 0
    w = random_weights(4, 64, 23)
    quant = ((-1.1, -1.05, -1.2), (2.2, 2.1, 2.4)) if quantized else None
    bits = None
    if exact:
        bits = np.zeros((mesh.positions.shape[0] + 31) // 32, np.uint32)
        bits[::3] = 0x5555AAAA
    with ActionSet(mesh, uv, tangents=tangents) as s, Policy(w, 4, 64) as p:
        mapping = s.mapping()
        wedge_checks(mesh, s.legal, mapping)
        rows = s.features(c, width=128)
        rows[:, 79] = 1.0  # architecture 4 carries preserve_uv there
        got = s.placements(p, c, preserve_uv=True, quantization=quant, exact_bits=bits)
        default = s.placements()
    y = so.policy_forward(rows, w, 64, 12)
    want = so.policy_placements(mesh.positions, mesh.normals, s.legal, mapping, y, supervise, quant, bits)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    plain = so.policy_placements(mesh.positions, mesh.normals, s.legal, mapping)
    assert np.array_equal(default.view(np.uint32), plain.view(np.uint32))
    if quantized:
        snapped = sum(not np.array_equal(so.grid_position(r[:3], *quant), r[:3]) for r in
                      so.policy_placements(mesh.positions, None, s.legal, mapping, y))
        assert snapped > 0  # the grid moved some placements


@pytest.mark.gpu
def test_gpu_zero_offset_places_at_the_midpoint():
    # A placement policy with zero weights and zero bias: offset 0, so every
    # position is float((from + to) * .5), unsnapped.
    from blitzremesher_amd import meshes
    from blitzremesher_amd.actions import ActionSet, Policy
    mesh = meshes.with_attributes(meshes.icosphere(2, 0.02, 3), 4)
    with ActionSet(mesh) as s, Policy(np.zeros(policy_weight_count(3, 64), np.float32), 3, 64) as p:
        got = s.placements(p, CONDITIONS)
    pos = mesh.positions
    mid = ((pos[s.legal[:, 0]] + pos[s.legal[:, 1]]).astype(np.float32).astype(np.float64) * 0.5).astype(np.float32)
    assert np.array_equal(got[:, :3], mid)
