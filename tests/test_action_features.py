"""Action features: the reference's vertex graph channels (neural::graph,
src/neural/graph.cpp:35-187) and action feature rows (ActionState::actions,
src/neural/action.cpp:371-450) computed on the device.

CPU: the committed fixture (tests/golden/action_features.npz, written by
tests/golden/make_action_feature_golden.py from the reference build in
oracle/_ref) equals what the reference build produces now.
GPU: the HIP channels and rows equal the fixture bit for bit on meshes with
seams, colours, tangents, two materials, boundaries and closed surfaces, and
the reference build on a larger closed mesh; the 128-wide rows carry the
second wedge's vertex channels; bad conditions and widths are rejected.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent))
from feature_cases import CONDITIONS, cases  # noqa: E402
from oracle_lib import Actions, available  # noqa: E402

from blitzremesher_amd import meshes  # noqa: E402
from blitzremesher_amd.audit import Mesh  # noqa: E402

FIXTURE = Path(__file__).resolve().parent / "golden" / "action_features.npz"
NAMES = sorted(cases())


def fixture():
    return np.load(FIXTURE)


def fixture_case(data, name):
    def get(key):
        k = f"{name}/{key}"
        return data[k] if k in data else None
    mesh = Mesh(get("positions"), get("indices"), get("materials"), None, get("normals"), get("colors"))
    return mesh, get("uv"), get("tangents")


def bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.skipif(not available("ref"), reason="reference build (oracle/_ref) not built")
@pytest.mark.parametrize("name", NAMES)
def test_fixture_matches_reference_build(name):
    data = fixture()
    mesh, uv, tangents = fixture_case(data, name)
    ref = Actions("ref", mesh, uv, tangents)
    assert np.array_equal(ref.legal, data[f"{name}/legal"])
    assert np.array_equal(bits(ref.ref_graph_features()), bits(data[f"{name}/graph"]))
    assert np.array_equal(bits(ref.ref_features(data["conditions"])), bits(data[f"{name}/rows"]))
    ref.close()


def test_fixture_covers_every_graph_flag():
    # Seams, boundaries and material junctions all occur, and stream
    # presence varies, so each channel's branch is exercised.
    data = fixture()
    g = np.concatenate([data[f"{n}/graph"] for n in NAMES])
    for channel in (19, 20, 21):
        assert (g[:, channel] == 1).any() and (g[:, channel] == 0).any(), channel
    for channel in (15, 16, 17, 18):
        assert set(np.unique(g[:, channel])) == {0.0, 1.0}, channel
    rows = np.concatenate([data[f"{n}/rows"] for n in NAMES])
    assert (rows[:, 75] == 1).any() and (rows[:, 75] == 0).any()  # one- and two-wedge actions
    assert (rows[:, 61] != 0).any() and (rows[:, 77] != 0).any()  # UV and colour deltas


# ----------------------------------------------------------------------- GPU

@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_graph_channels_match_reference(name):
    from blitzremesher_amd.actions import ActionSet
    data = fixture()
    mesh, uv, tangents = fixture_case(data, name)
    with ActionSet(mesh, uv, tangents=tangents) as s:
        assert np.array_equal(s.legal, data[f"{name}/legal"])
        got = s.vertex_features()
    want = data[f"{name}/graph"]
    bad = np.argwhere(bits(got) != bits(want))
    assert bad.size == 0, f"{len(bad)} channels differ, first (vertex, channel) {bad[:5].tolist()}"


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_action_rows_match_reference(name):
    from blitzremesher_amd.actions import ActionSet
    data = fixture()
    mesh, uv, tangents = fixture_case(data, name)
    with ActionSet(mesh, uv, tangents=tangents) as s:
        got = s.features(data["conditions"])
    want = data[f"{name}/rows"]
    assert got.shape == want.shape
    bad = np.argwhere(bits(got) != bits(want))
    assert bad.size == 0, f"{len(bad)} values differ, first (row, channel) {bad[:5].tolist()}"


@pytest.mark.gpu
@pytest.mark.skipif(not available("ref"), reason="reference build (oracle/_ref) not built")
def test_gpu_rows_match_reference_build_on_a_larger_mesh():
    # 2,562 vertices with normals and colours, and a split copy with seams,
    # UVs, tangents and materials: every row and channel bit for bit.
    from blitzremesher_amd.actions import ActionSet
    from feature_cases import seamed
    ico = meshes.with_attributes(meshes.icosphere(4, 0.02, 8), 10)
    for mesh, uv, tangents in ((ico, None, None), seamed(meshes.icosphere(3, 0.02, 9), 12)):
        ref = Actions("ref", mesh, uv, tangents)
        with ActionSet(mesh, uv, tangents=tangents) as s:
            assert np.array_equal(s.legal, ref.legal)
            assert np.array_equal(bits(s.vertex_features()), bits(ref.ref_graph_features()))
            assert np.array_equal(bits(s.features(CONDITIONS)), bits(ref.ref_features(CONDITIONS)))
        ref.close()


@pytest.mark.gpu
def test_gpu_placement_rows_carry_the_second_wedge():
    # The 128-wide rows (src/neural/action_gpu.cu:725-729): channels 0..79 as
    # the 80-wide rows; for two-wedge actions 80..103 / 104..127 are the
    # channels of the second source / target vertex, which share the first
    # pair's positions; zero otherwise.
    from blitzremesher_amd.actions import ActionSet
    data = fixture()
    mesh, uv, tangents = fixture_case(data, "seamed_grid")
    with ActionSet(mesh, uv, tangents=tangents) as s:
        wide = s.features(CONDITIONS, width=128)
        narrow = s.features(CONDITIONS)
        graph = s.vertex_features()
    assert np.array_equal(bits(wide[:, :80]), bits(narrow))
    two = wide[:, 75] == 1
    assert two.any() and (~two).any()
    assert not wide[~two, 80:].any()
    rows = {tuple(r) for r in bits(graph)}
    for r in wide[two]:
        assert tuple(bits(r[80:104])) in rows and tuple(bits(r[104:128])) in rows
        assert np.array_equal(r[80:83], r[0:3]) and np.array_equal(r[104:107], r[24:27])


@pytest.mark.gpu
def test_gpu_face_ratio_follows_the_input_count():
    from blitzremesher_amd.actions import ActionSet
    mesh = meshes.icosphere(1, 0.02, 5)
    with ActionSet(mesh) as s:
        own = s.features(CONDITIONS)
        half = s.features(CONDITIONS, input_triangles=2 * mesh.triangles)
    assert (own[:, 66] == 1).all()
    assert (half[:, 66] == np.float32(0.5)).all()
    assert np.array_equal(np.delete(own, 66, 1), np.delete(half, 66, 1))


@pytest.mark.gpu
def test_gpu_features_reject_bad_arguments():
    from blitzremesher_amd._lib import BlitzError
    from blitzremesher_amd.actions import ActionSet
    with ActionSet(meshes.icosphere(1, 0.02, 5)) as s:
        c = CONDITIONS.copy()
        c[3] = np.inf
        with pytest.raises(BlitzError):
            s.features(c)
        with pytest.raises(BlitzError):
            s.features(CONDITIONS, width=96)
