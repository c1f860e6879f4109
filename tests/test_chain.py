"""The LOD chain (blitz_hip_generate) against the reference's own chain
search: tests/golden/chain_golden.json holds what detail::generate_with_hooks
(src/chain.cpp:1175), built from the reference's CPU sources into
oracle/_ref, emits under the CPU hooks of oracle/ref_shim.cpp for the cases of
tests/chain_cases.py. The GPU chain must emit the same LODs (index buffers),
the same audit records of every level, the same schedule and the same
generation counters, including the four-gate audited collapses.
"""
from __future__ import annotations

import hashlib
import json
import math
from pathlib import Path

import numpy as np
import pytest

from blitzremesher_amd import Mesh, meshes
from blitzremesher_amd.chain import ChainOptions, ChainSettings, schedule
from chain_cases import chain_cases
from oracle_lib import HookSettings, Oracle, available, ref_generate_hooks

GOLDEN = Path(__file__).resolve().parent / "golden"
needs_ref = pytest.mark.skipif(not available("ref"), reason="reference CPU build (oracle/_ref) not present")
COUNTERS = ("status", "candidate_evaluations", "audit_evaluations", "rejected_gates", "legal_collapses",
            "rejected_collapses", "action_trials", "action_ranked", "decoded", "finalists", "fallback_levels")


def golden() -> dict:
    return {c["case"]: c for c in json.loads((GOLDEN / "chain_golden.json").read_text())["cases"]}


def golden_mesh(name: str) -> Mesh:
    z = np.load(GOLDEN / "chain_meshes.npz")
    return Mesh(z[f"{name}/positions"], z[f"{name}/indices"])


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.uint32).tobytes()).hexdigest()


def summary(lods) -> tuple:
    return [int(l.shape[0]) for l in lods], [sha(l) for l in lods]


def same_float(a, b) -> bool:
    return a == b or (isinstance(a, float) and isinstance(b, float) and math.isnan(a) and math.isnan(b))


def test_chain_case_meshes_match_fixtures():
    for name, (mesh, _, _) in chain_cases().items():
        g = golden_mesh(name)
        assert np.array_equal(mesh.positions, g.positions) and np.array_equal(mesh.indices, g.indices), name


def test_golden_chains_are_nontrivial():
    # The fixtures exercise reduction at every level, rejected trials and
    # several finalists, so a chain that stops early or skips a gate differs.
    for name, g in golden().items():
        c = g["counters"]
        assert c["status"] == 0 and c["legal_collapses"] > 0 and c["rejected_collapses"] > 0, name
        assert g["triangles"][-1] < g["triangles"][0] and c["finalists"] >= 2, name
        for i in range(1, len(g["triangles"])):
            assert g["source"][i][10] and g["adjacent"][i][10], (name, i)  # every level confirmed passed


@pytest.mark.parametrize("name", sorted(chain_cases()))
def test_host_schedule_matches_golden(name):
    mesh, settings, _ = chain_cases()[name]
    o = Oracle("oracle")
    center, radius = o.bounds(mesh)
    assert schedule(settings, center, radius) == golden()[name]["schedule"]


@needs_ref
def test_host_schedule_matches_reference_build():
    import ctypes as C
    rng = np.random.default_rng(17)
    lib = Oracle("ref").lib
    f = lib.ref_schedule
    f.restype = C.c_int
    for trial in range(200):
        levels = int(rng.integers(2, 33))
        base = float(rng.choice([0.0, rng.uniform(20, 4000)]))
        last = float(rng.uniform(1, 20))
        lod0 = float(rng.choice([0.0, rng.uniform(0.1, 5)]))
        xs = np.sort(rng.uniform(0, 1, int(rng.integers(0, 4))))
        pts = [(0.0, float(rng.uniform(0, 4)))] + [(float(x), float(rng.uniform(0, 4))) for x in xs] + \
              [(1.0, float(rng.uniform(0, 4)))]
        if len({p[0] for p in pts}) != len(pts):
            continue
        ppm, mpu, radius = float(rng.uniform(50, 900)), float(rng.uniform(0.2, 3)), float(rng.uniform(0.1, 4))
        if base == 0.0 and not (last <= 2 * radius * mpu * ppm <= 16384):
            continue
        s = ChainSettings(levels=levels, base_pixels=base, last_pixels=last, max_lod0_delta_px=lod0,
                          transition=tuple(pts), pixels_per_meter=ppm, meters_per_unit=mpu)
        hs = HookSettings(levels=levels, base_pixels=base, last_pixels=last, max_lod0_delta_px=lod0)
        tr = np.asarray(pts, np.float32).reshape(-1)
        hs.transition_count = len(pts)
        hs.transition = tr.ctypes.data_as(C.POINTER(C.c_float))
        want = np.zeros(3 * levels)
        center = (C.c_float * 3)(0.1, -0.2, 0.3)
        assert f(C.byref(hs), C.c_double(ppm), C.c_double(mpu), center, C.c_double(radius),
                 want.ctypes.data_as(C.POINTER(C.c_double))) == levels
        got = schedule(s, (0.1, -0.2, 0.3), radius)
        assert got == want.reshape(-1, 3).tolist(), (trial, s)


@needs_ref
def test_reference_build_reproduces_tiny_golden():
    mesh, settings, options = chain_cases()["ico1_tiny"]
    r = ref_generate_hooks(mesh, (settings, options))
    g = golden()["ico1_tiny"]
    assert summary(r["lods"]) == (g["triangles"], g["indices_sha256"])
    assert {k: r[k] for k in COUNTERS} == g["counters"]


def check_chain(result, g, name):
    got = summary([l.indices for l in result.lods])
    assert got == (g["triangles"], g["indices_sha256"]), (name, got[0], g["triangles"])
    assert [[l.screen_pixels, l.transition_limit, l.source_limit] for l in result.lods] == g["schedule"], name
    for i, l in enumerate(result.lods):
        for which, m in (("source", l.source), ("adjacent", l.adjacent)):
            want = g[which][i]
            assert all(same_float(a, b) for a, b in zip(m.as_tuple(), want)), (name, i, which, m, want)
    info = result.info
    counters = {k: (info[k] if k not in ("finalists",) else len(result.finalists)) for k in COUNTERS if k != "status"}
    counters["status"] = result.status
    assert counters == g["counters"], (name, counters, g["counters"])
    assert result.finalists == g["finalist_triangles"], name


@pytest.mark.gpu
@pytest.mark.parametrize("cache", [False, True])
@pytest.mark.parametrize("name", sorted(chain_cases()))
def test_gpu_chain_matches_reference_golden(name, cache):
    from blitzremesher_amd import Audit
    from blitzremesher_amd.chain import generate
    _, settings, options = chain_cases()[name]
    with Audit(0, cache_rasters=cache) as audit:
        r = generate(audit, golden_mesh(name), settings, options)
    print(name, cache, [l.triangles for l in r.lods], r.info["legal_collapses"], round(r.info["ms"], 1), "ms",
          flush=True)
    check_chain(r, golden()[name], name)


@pytest.mark.gpu
def test_gpu_chains_in_flight_match_reference_golden():
    """Several chains at once, one audit object per thread (as the batched
    generation runs them): each still emits the reference's chain."""
    import threading
    from concurrent.futures import ThreadPoolExecutor
    from blitzremesher_amd import Audit
    from blitzremesher_amd.chain import generate
    names = sorted(chain_cases()) * 2
    local = threading.local()
    owned = []
    lock = threading.Lock()

    def run(name):
        if not hasattr(local, "audit"):
            local.audit = Audit(0, cache_rasters=True)
            with lock:
                owned.append(local.audit)
        _, settings, options = chain_cases()[name]
        return generate(local.audit, golden_mesh(name), settings, options)

    try:
        with ThreadPoolExecutor(4) as pool:
            results = list(pool.map(run, names))
    finally:
        for audit in owned:
            audit.close()
    for name, r in zip(names, results):
        check_chain(r, golden()[name], name)


@pytest.mark.gpu
def test_gpu_chain_overdraw_tiebreak_keeps_the_counts():
    # Ties are broken only among finalists of equal per-level triangle counts
    # and storage, each confirmed by full audits first.
    from blitzremesher_amd import Audit
    from blitzremesher_amd.chain import generate
    for name in ("ico2_plane", "heightfield_plane"):
        _, settings, options = chain_cases()[name]
        options = ChainOptions(**{**options.__dict__, "overdraw_tiebreak": True})
        with Audit(0) as audit:
            r = generate(audit, golden_mesh(name), settings, options)
        g = golden()[name]
        assert [l.triangles for l in r.lods] == g["triangles"], name
        assert all(l.source.passed and l.adjacent.passed for l in r.lods[1:]), name


@pytest.mark.gpu
def test_gpu_chain_rejects_unsupported_options():
    from blitzremesher_amd import Audit, BlitzError
    from blitzremesher_amd.chain import generate
    mesh, settings, options = chain_cases()["ico1_tiny"]
    with Audit(0) as audit:
        for change, status in ((dict(origin=1), 6), (dict(output=0), 6), (dict(confirmation=0), 6),
                               (dict(ranking=2), 6), (dict(action_batch=0), 1), (dict(ranking=0), 1)):
            with pytest.raises(BlitzError) as e:
                generate(audit, mesh, settings, ChainOptions(**{**options.__dict__, **change}))
            assert e.value.status == status, change
        for change in (dict(levels=1), dict(beam_width=0), dict(search_supersample=3), dict(max_supersample=2)):
            with pytest.raises(BlitzError) as e:
                generate(audit, mesh, ChainSettings(**{**settings.__dict__, **change}), options)
            assert e.value.status == 1, change
