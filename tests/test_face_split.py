"""Coverage audits drawn from a reference/candidate face split.

When the reference and the candidate share a vertex set, their coverage
rasters are drawn from three face lists: faces of the candidate paired with
a reference face of the same ordered triple and sidedness (drawn once into
both masks), reference-only and candidate-only faces. The masks, and so every
Measurement field, must equal the oracle's audit of the whole meshes. Each
case is audited one view at a time (so a wrong bit in any view shows in that
view's changed area and distance) and as one multi-view audit, with the
split forced on, and the split must have been used.

Cases: a heightfield and its quadric-collapsed copy (reused vertices); an
icosphere with two materials of different sidedness whose candidate drops,
duplicates, rotates (same corners, a different triple: not shared) and
re-materials faces and is shuffled; and the same pair under a radius that
pushes faces out of frame (the first out-of-frame face in each mesh's own
order cuts its mask) and puts perspective cameras inside the mesh (behind
the camera).
"""
from __future__ import annotations

import numpy as np
import pytest

from blitzremesher_amd import EvalSettings, Mesh, meshes
from blitzremesher_amd.audit import bounds, reduce_reuse
from oracle_lib import Oracle


def mixed_pair(seed: int = 11):
    rng = np.random.default_rng(seed)
    ref = meshes.icosphere(3, 0.05, seed)
    f = ref.indices.copy()
    f = np.concatenate([f, f[:40]])  # duplicated reference faces
    mat = rng.integers(0, 2, len(f)).astype(np.uint16)
    ref = Mesh(ref.positions, f, mat, np.array([0, 1], np.uint8))
    keep = rng.random(len(f)) > 0.2
    cf, cm = f[keep].copy(), mat[keep].copy()
    rot = rng.random(len(cf)) < 0.1
    cf[rot] = np.roll(cf[rot], 1, axis=1)
    flip = rng.random(len(cf)) < 0.1
    cm[flip] ^= 1
    dup = rng.random(len(cf)) < 0.05
    cf, cm = np.concatenate([cf, cf[dup]]), np.concatenate([cm, cm[dup]])
    order = rng.permutation(len(cf))
    cand = Mesh(ref.positions, cf[order], cm[order], np.array([0, 1], np.uint8))
    return ref, cand


def heightfield_pair():
    ref = meshes.heightfield(48, 0.1, 2)
    return ref, reduce_reuse(ref, int(round(ref.triangles * 0.75)))


CASES = {
    "heightfield": (heightfield_pair, 1.0, 0, True),
    "mixed_sidedness": (mixed_pair, 1.0, 0, False),
    "out_of_frame_and_behind": (mixed_pair, 0.55, 2, False),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_split_audit_matches_oracle(name):
    from blitzremesher_amd import Audit
    make, shrink, persp, two = CASES[name]
    ref, cand = make()
    center, radius = bounds(ref)
    radius *= shrink
    o = Oracle("oracle")
    base = dict(profile=0, screen_size=40, supersample=2, max_supersample=8, limit=1e3, max_changed_area=1.0,
                force_two_sided=two)
    with Audit(0, face_split=2) as audit:
        dref, dcand = audit.upload(ref), audit.upload(cand)
        # One view per audit: each view's own changed area and distance.
        for seed in range(12):
            s = EvalSettings(**base, orthographic=1 if seed % 3 or not persp else 0,
                             perspective=0 if seed % 3 or not persp else 1, rotation_seed=0x5EED0000 + seed)
            got = audit.evaluate(dref, dcand, center, radius, s)
            assert audit.last_timing()["split_launches"] >= 1
            want = o.evaluate(ref, cand, center, radius, s)
            assert got.as_tuple() == want.as_tuple(), (name, seed, got, want)
        s = EvalSettings(**base, orthographic=16, perspective=persp, rotation_seed=0xB1172026)
        got = audit.evaluate(dref, dcand, center, radius, s)
        print(f"{name}: {got}", flush=True)
        assert audit.last_timing()["split_launches"] >= 1
        want = o.evaluate(ref, cand, center, radius, s)
        assert got.as_tuple() == want.as_tuple(), (name, got, want)
        if shrink < 1:
            assert got.coverage_upper == float("inf")  # clipped: faces left the frame


@pytest.mark.gpu
def test_gpu_split_and_whole_meshes_give_the_same_measurements():
    """The same audits with the split never used and always used."""
    from blitzremesher_amd import Audit
    ref, cand = heightfield_pair()
    center, radius = bounds(ref)
    s = EvalSettings(profile=0, screen_size=64, supersample=4, max_supersample=16, limit=0.75, orthographic=24,
                     perspective=0, rotation_seed=0xB1172026, force_two_sided=True)
    got = {}
    for mode in (1, 2):
        with Audit(0, face_split=mode) as audit:
            dref, dcand = audit.upload(ref), audit.upload(cand)
            got[mode] = [audit.evaluate(dref, dcand, center, radius, s).as_tuple() for _ in range(2)]
            assert (audit.last_timing()["split_launches"] >= 1) == (mode == 2)
    assert got[1] == got[2]


@pytest.mark.gpu
def test_gpu_autotuned_raster_shapes_give_the_same_measurements():
    """The raster launch shape measured on the first launch of a shape class
    (every candidate drawn into the same masks) changes no Measurement field;
    the second audit of the class reuses the choice."""
    from blitzremesher_amd import Audit
    ref, cand = heightfield_pair()
    center, radius = bounds(ref)
    # A raster side no other test uses, so this shape class is new here.
    s = EvalSettings(profile=0, screen_size=57, supersample=4, max_supersample=4, limit=1e3, orthographic=24,
                     perspective=0, rotation_seed=0xB1172026, force_two_sided=True)
    with Audit(0, autotune=False) as audit:
        dref, dcand = audit.upload(ref), audit.upload(cand)
        want = audit.evaluate(dref, dcand, center, radius, s).as_tuple()
        assert audit.last_timing()["tuned_launches"] == 0
    with Audit(0, autotune=True) as audit:
        dref, dcand = audit.upload(ref), audit.upload(cand)
        first = audit.evaluate(dref, dcand, center, radius, s).as_tuple()
        tuned = audit.last_timing()["tuned_launches"]
        second = audit.evaluate(dref, dcand, center, radius, s).as_tuple()
        assert audit.last_timing()["tuned_launches"] == 0
    assert tuned >= 1
    assert first == want and second == want
    assert want == Oracle("oracle").evaluate(ref, cand, center, radius, s).as_tuple()
