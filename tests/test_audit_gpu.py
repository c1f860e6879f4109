"""GPU parity of the HIP coverage audit against the CPU oracle.

Every comparison is exact: masks bit-for-bit, Measurement fields with ==.
Cases follow the reference's own audit tests (tests/precision.cpp:24-31 and its
main(), tests/contracts.cpp:131-190, tests/neural.cpp:614-639).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from blitzremesher_amd import Audit, EvalSettings, Mesh, meshes
from oracle_lib import Oracle, mask_to_words, words_to_mask

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def audit():
    a = Audit(0)
    yield a
    a.close()


@pytest.fixture(scope="module")
def oracle():
    return Oracle("oracle")


def assert_same_measurement(gpu, cpu, label=""):
    assert gpu.as_tuple() == cpu.as_tuple(), f"{label}\n gpu={gpu}\n cpu={cpu}"


def raster_pair(audit, oracle, mesh, center, radius, cam, screen, ss, two):
    d = audit.upload(mesh)
    g = audit.raster_coverage(d, center, radius, cam, screen, ss, two)
    c = oracle.raster(mesh, center, radius, cam, screen, ss, two)
    d.release()
    return g, c


def test_raster_masks_grid_views(audit, oracle):
    # tests/precision.cpp: grid(4), screen 17.25, views {8,4,seed}, ss 1..8.
    m = meshes.grid(4)
    center, radius = oracle.bounds(m)
    d = audit.upload(m)
    checked = 0
    for seed in (17, 9381):
        cams = oracle.cameras(center, radius, 17.25, 8, 4, seed)
        for ss in (1, 2, 4, 8):
            for cam in cams:
                for two in (False, True):
                    g = audit.raster_coverage(d, center, radius, cam, 17.25, ss, two)
                    c = oracle.raster(m, center, radius, cam, 17.25, ss, two)
                    assert g[1:] == c[1:]
                    assert np.array_equal(g[0], c[0]), (seed, ss, two)
                    checked += 1
    assert checked == 2 * 4 * 12 * 2


def test_raster_cull_degenerate_clip_behind(audit, oracle):
    # Culling, zero-area faces, subpixel geometry, clipping and behind-camera
    # vertices (tests/precision.cpp main()).
    m = meshes.grid(4)
    center, radius = oracle.bounds(m)
    cam = oracle.cameras(center, radius, 17.25, 1, 1, 99)[-1]
    idx = np.concatenate([m.indices, np.zeros((1, 3), np.uint32)])
    pos = m.positions.copy()
    pos[:, 2] += np.float32(1e-5)
    one_sided = Mesh(pos, idx, double_sided=np.array([0], np.uint8))
    g, c = raster_pair(audit, oracle, one_sided, center, radius, cam, 17.25, 2, False)
    assert g[1:] == c[1:] and np.array_equal(g[0], c[0])
    pos2 = pos.copy()
    pos2[:, 0] += np.float32(100)
    moved = Mesh(pos2, idx, double_sided=np.array([0], np.uint8))
    g, c = raster_pair(audit, oracle, moved, center, radius, cam, 17.25, 2, True)
    assert c[2] and g[1:] == c[1:] and np.array_equal(g[0], c[0])
    cam.distance = -100.0
    g, c = raster_pair(audit, oracle, moved, center, radius, cam, 17.25, 2, True)
    assert g[1:] == c[1:] and np.array_equal(g[0], c[0])


def test_raster_partial_out_of_frame_order(audit, oracle):
    # The reference stops at the first out-of-frame face; faces before it stay.
    m = meshes.icosphere(2, 0.0)
    center, radius = oracle.bounds(m)
    pos = m.positions.copy()
    far = m.indices[len(m.indices) // 2]
    pos[far[0]] += np.float32(5.0)
    moved = Mesh(pos, m.indices)
    for cam in oracle.cameras(center, radius, 30, 3, 2, 5):
        g, c = raster_pair(audit, oracle, moved, center, radius, cam, 30, 2, False)
        assert g[1:] == c[1:] and np.array_equal(g[0], c[0])


@pytest.mark.parametrize("screen", [23, 24, 25, 57, 58, 59, 121])
def test_raster_tile_straddling_screens(audit, oracle, screen):
    m = meshes.torus(1.0, 0.3, 40, 16, 0.01, 3)
    center, radius = oracle.bounds(m)
    d = audit.upload(m)
    for cam in oracle.cameras(center, radius, screen, 3, 1, 0xB1172026):
        g = audit.raster_coverage(d, center, radius, cam, screen, 2, False)
        c = oracle.raster(m, center, radius, cam, screen, 2, False)
        assert g[1:] == c[1:] and np.array_equal(g[0], c[0])


def test_raster_large_triangles(audit, oracle):
    # Few, large faces exercise the wavefront tile walk.
    m = meshes.icosphere(0, 0.0)
    center, radius = oracle.bounds(m)
    d = audit.upload(m)
    for cam in oracle.cameras(center, radius, 100, 4, 2, 3):
        for ss in (1, 4):
            g = audit.raster_coverage(d, center, radius, cam, 100, ss, False)
            c = oracle.raster(m, center, radius, cam, 100, ss, False)
            assert g[1:] == c[1:] and np.array_equal(g[0], c[0])


def test_raster_nonfinite_positions(audit, oracle):
    # NaN and infinite coordinates: out-of-frame stops, finite boxes with NaN
    # edge tests, both rasters exactly as the oracle's.
    from golden_data import nonfinite_meshes
    for name, (m, c, r) in nonfinite_meshes().items():
        d = audit.upload(m)
        for cam in oracle.cameras(c, r, 24, 3, 2, 77):
            for ss, two in ((1, False), (4, True)):
                g = audit.raster_coverage(d, c, r, cam, 24, ss, two)
                o = oracle.raster(m, c, r, cam, 24, ss, two)
                assert g[1:] == o[1:] and np.array_equal(g[0], o[0]), (name, ss, two)
        d.release()


def test_coverage_distance_random_masks(audit, oracle):
    # tests/contracts.cpp:115-121 and tests/precision.cpp random masks (square here).
    rng = np.random.default_rng(871)
    for trial in range(40):
        size = int(rng.integers(1, 40))
        a = rng.random((size, size)) < 0.2
        b = rng.random((size, size)) < 1 / 7
        wa, wb = mask_to_words(a), mask_to_words(b)
        for ss in (1, 2, 8):
            assert audit.coverage_distance(wa, wb, size, ss) == oracle.coverage_distance(wa, wb, size, ss)


def test_coverage_distance_special(audit, oracle):
    empty = mask_to_words(np.zeros((1, 1), bool))
    one = mask_to_words(np.ones((1, 1), bool))
    assert audit.coverage_distance(empty, empty, 1, 1) == 0
    assert math.isinf(audit.coverage_distance(empty, one, 1, 1))
    # Two single samples at offset (dx, dy): sqrt(dx^2 + dy^2) / ss.
    a = np.zeros((64, 64), bool)
    b = np.zeros((64, 64), bool)
    a[3, 5] = True
    b[40, 61] = True
    assert audit.coverage_distance(mask_to_words(a), mask_to_words(b), 64, 4) == math.sqrt(37 ** 2 + 56 ** 2) / 4


def test_coverage_distance_large_rounding(audit, oracle):
    # Squared distances above 2^24 round through float exactly as the reference.
    size = 4200
    a = np.zeros((size, size), bool)
    b = np.zeros((size, size), bool)
    a[0, 0] = True
    a[size - 1, size - 1] = True
    b[0, 0] = True
    b[size - 1, 17] = True
    wa, wb = mask_to_words(a), mask_to_words(b)
    assert audit.coverage_distance(wa, wb, size, 1) == oracle.coverage_distance(wa, wb, size, 1)


def eval_settings(**kw):
    base = dict(profile=0, screen_size=40, supersample=2, max_supersample=8, orthographic=6, perspective=2,
                rotation_seed=0xB1172026, limit=1.0)
    base.update(kw)
    return EvalSettings(**base)


def check_evaluate(audit, oracle, ref, cand, center, radius, s, label=""):
    g = audit.evaluate(ref, cand, center, radius, s)
    c = oracle.evaluate(ref, cand, center, radius, s)
    assert_same_measurement(g, c, label)
    return g


@pytest.mark.parametrize("limit", [0.25, 1.0, 3.0, 1000.0])
def test_evaluate_icosphere_jitter(audit, oracle, limit):
    ico = meshes.icosphere(3, 0.02, 1)
    center, radius = oracle.bounds(ico)
    cand = meshes.jitter(ico, 0.004, 5, 0.3)
    check_evaluate(audit, oracle, ico, cand, center, radius, eval_settings(limit=limit), f"limit={limit}")


def test_evaluate_refinement_ladder(audit, oracle):
    ico = meshes.icosphere(3, 0.02, 1)
    center, radius = oracle.bounds(ico)
    cand = meshes.jitter(ico, 0.002, 9, 0.5)
    for ss, mx, limit in [(1, 8, 0.6), (2, 16, 0.8), (3, 12, 1.0), (4, 4, 1.0), (1, 32, 0.5)]:
        s = eval_settings(supersample=ss, max_supersample=mx, limit=limit, screen_size=33)
        check_evaluate(audit, oracle, ico, cand, center, radius, s, f"ss={ss} max={mx} limit={limit}")


def test_evaluate_contract_cases(audit, oracle):
    # tests/contracts.cpp:131-134: identical passes, +2 translation fails,
    # a 4096 px screen is resource limited.
    m = meshes.grid(8)
    center, radius = oracle.bounds(m)
    s = EvalSettings(screen_size=24, orthographic=8, perspective=2, rotation_seed=1, supersample=4,
                     max_supersample=8, profile=0)
    g = check_evaluate(audit, oracle, m, m, center, radius, s, "identical")
    assert g.passed and g.views_evaluated == 0
    moved = Mesh(m.positions + np.array([2, 0, 0], np.float32), m.indices, double_sided=m.double_sided)
    g = check_evaluate(audit, oracle, m, moved, center, radius, s, "moved")
    assert not g.passed
    s.screen_size, s.supersample, s.max_supersample = 4096, 8, 8
    g = check_evaluate(audit, oracle, m, moved, center, radius, s, "limited")
    assert g.resource_limited and not g.passed


def test_evaluate_changed_area_tiles(audit, oracle):
    # tests/contracts.cpp:136-190: four tiles minus one, changed-area gate.
    tiles = meshes.four_tiles()
    center, radius = oracle.bounds(tiles)
    minus_first = Mesh(tiles.positions, tiles.indices[2:], double_sided=tiles.double_sided)
    minus_last = Mesh(tiles.positions, tiles.indices[:6], double_sided=tiles.double_sided)
    empty = Mesh(tiles.positions, np.zeros((0, 3), np.uint32), double_sided=tiles.double_sided)
    other_empty = Mesh(tiles.positions + np.array([.125, 0, 0], np.float32), np.zeros((0, 3), np.uint32),
                       double_sided=tiles.double_sided)
    for views in (1, 4):
        for mca in (1.0, 0.35, 0.1, 0.0):
            s = EvalSettings(profile=0, orthographic=views, perspective=0, rotation_seed=460796, screen_size=64,
                             supersample=4, max_supersample=4, force_two_sided=True, limit=1000,
                             max_changed_area=mca, normal_weight=0, color_weight=0, material_weight=0)
            for ref, cand in [(tiles, minus_first), (tiles, minus_last), (minus_first, minus_last), (tiles, tiles),
                              (empty, other_empty), (tiles, empty), (empty, tiles)]:
                check_evaluate(audit, oracle, ref, cand, center, radius, s, f"views={views} mca={mca}")


@pytest.mark.parametrize("screen", [23, 24, 25, 57, 58, 59, 121])
def test_evaluate_cut_patterns(audit, oracle, screen):
    # tests/neural.cpp:614-639: cut patterns straddling tile boundaries.
    m = meshes.grid(24)
    center, radius = oracle.bounds(m)
    cut = Mesh(m.positions, m.indices[: len(m.indices) * 2 // 3], double_sided=m.double_sided)
    s = eval_settings(screen_size=screen, orthographic=3, perspective=1, supersample=2, max_supersample=4,
                      limit=2.0)
    check_evaluate(audit, oracle, m, cut, center, radius, s, f"screen={screen}")


def test_evaluate_perspective_and_conservative(audit, oracle):
    t = meshes.torus(1.0, 0.3, 60, 24, 0.01, 3)
    center, radius = oracle.bounds(t)
    cand = meshes.drop_faces(t, 0.02, 4)
    for conservative in (False, True):
        s = eval_settings(orthographic=2, perspective=6, conservative_screen=conservative, limit=1.5)
        check_evaluate(audit, oracle, t, cand, center, radius, s, f"conservative={conservative}")


def test_evaluate_cancellation(audit, oracle):
    ico = meshes.icosphere(2, 0.02, 1)
    center, radius = oracle.bounds(ico)
    cand = meshes.jitter(ico, 0.01, 5)
    for stop_at in (0, 1, 3):
        calls = {"g": 0, "c": 0}

        def make(key):
            def cancelled():
                calls[key] += 1
                return calls[key] > stop_at
            return cancelled
        sg = eval_settings(limit=1000.0, cancelled=make("g"))
        sc = eval_settings(limit=1000.0, cancelled=make("c"))
        g = audit.evaluate(ico, cand, center, radius, sg)
        c = oracle.evaluate(ico, cand, center, radius, sc)
        assert_same_measurement(g, c, f"stop_at={stop_at}")
        assert g.cancelled and calls["g"] == calls["c"]


def test_evaluate_batch_matches_single(audit, oracle):
    ico = meshes.icosphere(3, 0.02, 1)
    center, radius = oracle.bounds(ico)
    ref = audit.upload(ico)
    cands = [meshes.jitter(ico, 0.001 * (k + 1), 20 + k, 0.3) for k in range(5)] + [ico]
    dev = [audit.upload(c) for c in cands]
    s = eval_settings(limit=1.0)
    batch = audit.evaluate_batch(ref, dev, center, radius, s)
    for k, (c, b) in enumerate(zip(cands, batch)):
        single = audit.evaluate(ref, dev[k], center, radius, s)
        assert b.as_tuple() == single.as_tuple()
        assert_same_measurement(b, oracle.evaluate(ico, c, center, radius, s), f"candidate {k}")


def test_unsupported_patch_profile_is_loud(audit, oracle):
    # Bound references (patch audits) implement the coverage profile only.
    m = meshes.grid(4)
    center, radius = oracle.bounds(m)
    with pytest.raises(Exception) as e:
        audit.bind(m, center, radius, EvalSettings(profile=1, screen_size=16, supersample=1, max_supersample=1,
                                                   orthographic=1, perspective=0))
    assert "UNAVAILABLE" in str(e.value)


def test_gpu_rasters_match_golden(audit):
    # Mask digests written by the reference's own CPU sources.
    from golden_data import audit_golden, camera_of, digest, golden_meshes
    ms = golden_meshes()
    dev = {}
    for case in audit_golden()["rasters"]:
        if case["mesh"] not in dev:
            dev[case["mesh"]] = audit.upload(ms[case["mesh"]])
        r = audit.raster_coverage(dev[case["mesh"]], case["center"], case["radius"], camera_of(case["camera"]),
                                  case["screen"], case["ss"], case["two_sided"])
        if case.get("limited"):
            assert r is None
            continue
        words, size, clipped = r
        assert (size, clipped, digest(words)) == (case["size"], case["clipped"], case["sha256"]), case


def test_gpu_distances_match_golden(audit):
    from golden_data import audit_golden, unpack_mask
    for case in audit_golden()["distances"]:
        a = unpack_mask(case["a"], case["size"])
        b = unpack_mask(case["b"], case["size"])
        assert audit.coverage_distance(mask_to_words(a), mask_to_words(b), case["size"], case["ss"]) == case["value"]


def test_gpu_evaluations_match_golden(audit):
    from golden_data import audit_golden, golden_meshes, measurement_of, settings_of
    ms = golden_meshes()
    dev = {name: audit.upload(m) for name, m in ms.items()}
    for case in audit_golden()["evaluations"]:
        got = audit.evaluate(dev[case["reference"]], dev[case["candidate"]], case["center"], case["radius"],
                             settings_of(case["settings"]))
        assert got == measurement_of(case["measurement"]), case["case"]


def test_reference_mask_cache_keeps_results(oracle):
    # Audits against one uploaded reference with its masks kept across calls
    # equal fresh audits: refinement, first failures, perspective views, a
    # change of bounds or screen (new cache context), batches.
    ico = meshes.icosphere(3, 0.02, 1)
    center, radius = oracle.bounds(ico)
    cands = [meshes.jitter(ico, 0.002 * (k + 1), 40 + k, 0.4) for k in range(4)]
    settings = [eval_settings(limit=1.0), eval_settings(limit=0.6), eval_settings(limit=3.0)]
    with Audit(0, cache_rasters=True) as cached, Audit(0) as fresh:
        rc, rf = cached.upload(ico), fresh.upload(ico)
        for s in settings + settings:
            for c in cands:
                got = cached.evaluate(rc, cached.upload(c), center, radius, s)
                want = fresh.evaluate(rf, fresh.upload(c), center, radius, s)
                assert got.as_tuple() == want.as_tuple()
        s = eval_settings(limit=1.0)
        moved = (center + np.float32(0.01), radius * 1.1)
        assert cached.evaluate(rc, cached.upload(cands[0]), *moved, s).as_tuple() == \
            fresh.evaluate(rf, fresh.upload(cands[0]), *moved, s).as_tuple()
        batch = cached.evaluate_batch(rc, [cached.upload(c) for c in cands], center, radius, s)
        for c, m in zip(cands, batch):
            assert m.as_tuple() == fresh.evaluate(rf, fresh.upload(c), center, radius, s).as_tuple()
