"""Appearance audit (normals and attributes profiles).

CPU: the oracle's attribute raster and attributed evaluation are pinned to
golden fixtures written by the reference's own CPU sources (per-stream
digests of attribute rasters, full Measurements of profiles 1 and 2), to
that build itself on fresh seeded inputs (oracle/_ref, when present), and
metric_acos to its closed forms and to libm within two ulps.

GPU: the HIP attribute raster equals the golden digests and the oracle byte
for byte (covered, visible, stored depth, normal, colour, material); HIP
Measurements of profiles 1 and 2 equal the golden ones and the oracle's
field for field, over refinement, first failures, clipped and behind-camera
views, materials, two-sided faces, zero weights, conservative screens and
batches.
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


# ----------------------------------------------------------------------- GPU

def _gpu_audit():
    from blitzremesher_amd import Audit
    return Audit(0)


def _same_pixels(got, want, label):
    g, w = got[0], want[0]
    assert got[1:] == want[1:], label
    for k in w:
        assert np.array_equal(g[k].view(np.uint8), w[k].view(np.uint8)), (label, k)


@pytest.mark.gpu
def test_gpu_pixels_match_golden():
    ms = golden_meshes()
    cases = audit_golden()["pixels"]
    with _gpu_audit() as audit:
        dev = {}
        for case in cases:
            if case["mesh"] not in dev:
                dev[case["mesh"]] = audit.upload(ms[case["mesh"]])
            got = audit.raster_pixels(dev[case["mesh"]], case["center"], case["radius"], camera_of(case["camera"]),
                                      case["screen"], case["ss"], case["two_sided"])
            if case.get("limited"):
                assert got is None
                continue
            streams, size, clipped = got
            assert (size, clipped) == (case["size"], case["clipped"])
            assert {k: pixel_digest(v) for k, v in streams.items()} == case["sha256"], case


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(attributed_cases()))
def test_gpu_pixels_match_oracle(case):
    mesh = attributed_cases()[case]
    o = Oracle("oracle")
    center, radius = o.bounds(mesh)
    with _gpu_audit() as audit:
        d = audit.upload(mesh)
        for screen in (24, 61):
            for cam in o.cameras(center, radius, screen, 5, 3, 77):
                for ss in (1, 3):
                    for two in (False, True):
                        got = audit.raster_pixels(d, center, radius, cam, screen, ss, two)
                        want = o.raster_pixels(mesh, center, radius, cam, screen, ss, two)
                        _same_pixels(got, want, (case, screen, ss, two))


@pytest.mark.gpu
def test_gpu_pixels_dense_overlap_and_clipping():
    # Many faces per tile (a fine sphere at a small screen), depth ties
    # between coincident duplicated faces, and perspective views whose
    # camera sits inside or behind the mesh (clipped faces).
    o = Oracle("oracle")
    ico = meshes.with_attributes(meshes.icosphere(4, 0.0, 1), 3)
    dup = Mesh(ico.positions, np.concatenate([ico.indices, ico.indices[::-1], ico.indices[:, ::-1]]),
               normals=ico.normals, colors=ico.colors)
    with _gpu_audit() as audit:
        for mesh in (ico, dup):
            center, radius = o.bounds(mesh)
            d = audit.upload(mesh)
            for cam in o.cameras(center, radius, 12, 3, 3, 5):
                got = audit.raster_pixels(d, center, radius, cam, 12, 2, True)
                want = o.raster_pixels(mesh, center, radius, cam, 12, 2, True)
                _same_pixels(got, want, "dense")
            persp = o.cameras(center, radius, 12, 0, 2, 9)
            for dist in (0.5 * radius, -radius):
                cam = persp[0]
                cam.distance = dist
                got = audit.raster_pixels(d, center, radius, cam, 12, 2, True)
                want = o.raster_pixels(mesh, center, radius, cam, 12, 2, True)
                assert want[2] is True
                _same_pixels(got, want, ("clipped", dist))


@pytest.mark.gpu
def test_gpu_attributed_evaluations_match_golden():
    ms = golden_meshes()
    cases = audit_golden()["attributed"]
    with _gpu_audit() as audit:
        for case in cases:
            got = audit.evaluate(ms[case["reference"]], ms[case["candidate"]], case["center"], case["radius"],
                                 settings_of(case["settings"]))
            assert got == measurement_of(case["measurement"]), (case["case"], case["reference"],
                                                                case["candidate"], got)


@pytest.mark.gpu
@pytest.mark.parametrize("profile", [1, 2])
def test_gpu_attributed_matches_oracle(profile):
    cases = attributed_cases()
    o = Oracle("oracle")
    pairs = [("icosphere", "icosphere_moved"), ("heightfield", "materials"), ("no_streams", "icosphere"),
             ("icosphere", "icosphere"), ("icosphere_moved", "icosphere")]
    with _gpu_audit() as audit:
        for ref_name, cand_name in pairs:
            ref, cand = cases[ref_name], cases[cand_name]
            center, radius = o.bounds(ref)
            for limit in (0.3, 1.0, 2.5, 40.0):
                for ss, mx, persp in ((1, 4, 3), (2, 8, 0), (3, 3, 5)):
                    s = EvalSettings(profile=profile, screen_size=28, supersample=ss, max_supersample=mx,
                                     orthographic=5, perspective=persp, limit=limit, max_changed_area=1.0)
                    got = audit.evaluate(ref, cand, center, radius, s)
                    want = o.evaluate(ref, cand, center, radius, s)
                    assert got == want, (profile, ref_name, cand_name, limit, ss, got, want)


@pytest.mark.gpu
def test_gpu_attributed_weights_and_modes():
    cases = attributed_cases()
    o = Oracle("oracle")
    ref, cand = cases["heightfield"], cases["materials"]
    center, radius = o.bounds(ref)
    settings = [
        dict(profile=2, normal_weight=0.0, color_weight=0.0, material_weight=0.0),   # zero weights
        dict(profile=2, normal_weight=0.0, color_weight=0.0, material_weight=2.0),   # materials only
        dict(profile=2, normal_weight=0.0, color_weight=9.0, material_weight=0.0),   # colours only
        dict(profile=1, normal_weight=40.0),
        dict(profile=1, conservative_screen=True),                                    # coverage rule
        dict(profile=2, force_two_sided=True, max_changed_area=0.02),
    ]
    with _gpu_audit() as audit:
        for extra in settings:
            for limit in (0.8, 3.0, 100.0):
                s = EvalSettings(**{**dict(screen_size=30, supersample=2, max_supersample=4, orthographic=6,
                                           perspective=2, limit=limit), **extra})
                got = audit.evaluate(ref, cand, center, radius, s)
                want = o.evaluate(ref, cand, center, radius, s)
                assert got == want, (extra, limit, got, want)


@pytest.mark.gpu
def test_gpu_attributed_batch_matches_single():
    o = Oracle("oracle")
    base = meshes.with_attributes(meshes.icosphere(3, 0.02, 1), 4)
    cands = [meshes.with_attributes(meshes.jitter(meshes.icosphere(3, 0.02, 1), 0.002 * (k + 1), 30 + k, 0.3),
                                    4 + (k % 2)) for k in range(4)] + [base]
    center, radius = o.bounds(base)
    s = EvalSettings(profile=2, screen_size=40, supersample=2, max_supersample=8, orthographic=8, perspective=4,
                     limit=2.0)
    with _gpu_audit() as audit:
        ref = audit.upload(base)
        dev = [audit.upload(c) for c in cands]
        batch = audit.evaluate_batch(ref, dev, center, radius, s)
        timing = audit.last_timing()
        assert timing["passes"] >= 1 and timing["appearance_ms"] > 0
        for k, (c, b) in enumerate(zip(cands, batch)):
            assert b == audit.evaluate(ref, dev[k], center, radius, s)
            assert b == o.evaluate(base, c, center, radius, s), k
        assert any(b.normal_degrees > 0 for b in batch)
