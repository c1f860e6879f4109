"""Appearance audit (normals and attributes profiles).

CPU: the oracle's attribute raster and attributed evaluation are pinned to
golden fixtures written by the reference's own CPU sources (per-stream
digests of attribute rasters, full Measurements of profiles 1 and 2), to
that build itself on fresh seeded inputs (oracle/_ref, when present), and
metric_acos to its closed forms and to libm within two ulps.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from blitzremesher_amd import EvalSettings, Mesh, meshes
from golden_data import audit_golden, camera_of, golden_meshes, measurement_of, pixel_digest, settings_of
from oracle_lib import Oracle, available

needs_ref = pytest.mark.skipif(not available("ref"), reason="reference CPU build (oracle/_ref) not present")


def attributed_cases():
    ico = meshes.with_attributes(meshes.icosphere(2, 0.02, 1), 7)
    ico_moved = meshes.with_attributes(meshes.jitter(meshes.icosphere(2, 0.02, 1), 0.01, 3, 0.4), 7)
    hf = meshes.with_attributes(meshes.heightfield(20, 0.1, 2), 9, colors=True)
    hf_mat = Mesh(hf.positions, hf.indices, (np.arange(hf.triangles) % 3).astype(np.uint16),
                  np.array([0, 1, 0], np.uint8), hf.normals, hf.colors)
    plain = meshes.icosphere(2, 0.02, 1)
    return {"icosphere": ico, "icosphere_moved": ico_moved, "heightfield": hf, "materials": hf_mat,
            "no_streams": plain}


def test_pixels_match_golden():
    o = Oracle("oracle")
    ms = golden_meshes()
    cases = audit_golden()["pixels"]
    assert len(cases) >= 50
    for case in cases:
        got = o.raster_pixels(ms[case["mesh"]], case["center"], case["radius"], camera_of(case["camera"]),
                              case["screen"], case["ss"], case["two_sided"])
        if case.get("limited"):
            assert got is None
            continue
        streams, size, clipped = got
        assert (size, clipped) == (case["size"], case["clipped"])
        assert {k: pixel_digest(v) for k, v in streams.items()} == case["sha256"], case


def test_attributed_evaluations_match_golden():
    o = Oracle("oracle")
    ms = golden_meshes()
    cases = audit_golden()["attributed"]
    assert len(cases) >= 40 and any(c["measurement"]["normal_degrees"] > 0 for c in cases)
    for case in cases:
        got = o.evaluate(ms[case["reference"]], ms[case["candidate"]], case["center"], case["radius"],
                         settings_of(case["settings"]))
        assert got == measurement_of(case["measurement"]), case["case"]


def test_metric_acos_closed_forms():
    o = Oracle("oracle")
    assert o.metric_acos(1.0) == 0.0
    assert abs(o.metric_acos(0.0) - math.pi / 2) < 1e-16
    assert abs(o.metric_acos(-1.0) - math.pi) < 1e-15
    rng = np.random.default_rng(0)
    for x in np.concatenate([rng.uniform(-1, 1, 2000), [0.5, -0.5, 0.5000001, -0.4999999, 1e-20, -1e-20]]):
        got, want = o.metric_acos(float(x)), math.acos(float(x))
        assert abs(got - want) <= 2 * math.ulp(want), (x, got, want)


@needs_ref
@pytest.mark.parametrize("case", sorted(attributed_cases()))
def test_oracle_pixels_match_reference_build(case):
    mesh = attributed_cases()[case]
    o, r = Oracle("oracle"), Oracle("ref")
    center, radius = o.bounds(mesh)
    for cam in o.cameras(center, radius, 24, 5, 3, 77):
        for ss in (1, 3):
            for two in (False, True):
                a = o.raster_pixels(mesh, center, radius, cam, 24, ss, two)
                b = r.raster_pixels(mesh, center, radius, cam, 24, ss, two)
                assert a[1:] == b[1:]
                for k in a[0]:
                    assert np.array_equal(a[0][k].view(np.uint8), b[0][k].view(np.uint8)), (case, k, ss, two)


@needs_ref
@pytest.mark.parametrize("profile", [1, 2])
def test_oracle_attributed_evaluate_matches_reference_build(profile):
    cases = attributed_cases()
    o, r = Oracle("oracle"), Oracle("ref")
    pairs = [("icosphere", "icosphere_moved"), ("heightfield", "materials"), ("no_streams", "icosphere"),
             ("icosphere", "icosphere")]
    for ref_name, cand_name in pairs:
        ref, cand = cases[ref_name], cases[cand_name]
        center, radius = o.bounds(ref)
        for limit in (0.5, 2.0, 40.0):
            for ss, mx in ((1, 4), (2, 2)):
                s = EvalSettings(profile=profile, screen_size=20, supersample=ss, max_supersample=mx, orthographic=6,
                                 perspective=3, limit=limit, max_changed_area=1.0)
                got = o.evaluate(ref, cand, center, radius, s)
                want = r.evaluate(ref, cand, center, radius, s)
                assert got.as_tuple() == want.as_tuple(), (ref_name, cand_name, limit, ss, got, want)
