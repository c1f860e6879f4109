"""Seeded randomized parity of full audits: the HIP audit against the CPU
oracle (a restatement of src/evaluate.cpp, itself pinned to the golden
vectors and the reference build) over configurations drawn from the space the
reference's tests span — coverage, normals and attributes profiles,
orthographic and perspective views, supersampling ladders, limits on both
sides of the measured error, one- and two-sided faces, clipped views and
reduced or jittered candidates. Every Measurement field must be equal."""
from __future__ import annotations

import numpy as np
import pytest

from blitzremesher_amd import Audit, EvalSettings, Mesh, meshes
from blitzremesher_amd.audit import reduce_reuse
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def _pair(rng: np.random.Generator, o: Oracle):
    kind = int(rng.integers(0, 5))
    seed = int(rng.integers(1, 1000))
    if kind == 0:
        ref = meshes.icosphere(int(rng.integers(1, 4)), 0.03, seed)
    elif kind == 1:
        ref = meshes.heightfield(int(rng.integers(6, 24)), float(rng.uniform(0.05, 0.4)), seed)
    elif kind == 2:
        ref = meshes.torus(1.0, float(rng.uniform(0.2, 0.45)), int(rng.integers(8, 24)), int(rng.integers(5, 12)),
                           0.01, seed)
    elif kind == 3:
        ref = meshes.grid(int(rng.integers(3, 10)))
    else:
        ref = meshes.four_tiles()
    if rng.random() < 0.5:
        ref = meshes.with_attributes(ref, seed, colors=bool(rng.random() < 0.7))
    mode = int(rng.integers(0, 3))
    if mode == 0 and ref.triangles > 8:
        red = reduce_reuse(ref, max(2, int(ref.triangles * rng.uniform(0.5, 0.95))))
        cand = Mesh(red.positions, red.indices, normals=ref.normals, colors=ref.colors)
    elif mode == 1:
        cand = meshes.jitter(ref, float(rng.uniform(0.001, 0.03)), seed + 1, float(rng.uniform(0.1, 1.0)))
    else:
        cand = ref
    return ref, cand


def test_random_audits_match_oracle():
    rng = np.random.default_rng(20261204)
    o = Oracle("oracle")
    checked = failed = 0
    with Audit(0) as audit:
        for case in range(120):
            ref, cand = _pair(rng, o)
            center, radius = o.bounds(ref)
            ss = int(rng.choice([1, 2, 3, 4]))
            s = EvalSettings(profile=int(rng.integers(0, 3)), screen_size=float(rng.uniform(12, 70)), supersample=ss,
                             max_supersample=int(ss * rng.choice([1, 2, 4])), orthographic=int(rng.integers(1, 7)),
                             perspective=int(rng.integers(0, 4)), rotation_seed=int(rng.integers(0, 2**32)),
                             limit=float(rng.choice([0.2, 0.7, 1.5, 4.0, 1e3])),
                             max_changed_area=float(rng.choice([1.0, 0.05])),
                             force_two_sided=bool(rng.random() < 0.4))
            got = audit.evaluate(ref, cand, center, radius, s)
            want = o.evaluate(ref, cand, center, radius, s)
            assert got == want, (case, s, got, want)
            checked += 1
            failed += not want.passed
    # Both verdicts occur, so both sides of every limit are exercised.
    assert checked == 120 and 10 < failed < 110
