"""Benchmark of the HIP coverage audit on the raster-audit workload.

Workload (one step): a full coverage audit of a seeded 512 x 512 fBm
heightfield (262,144 V / 522,242 F) against its 25 %-collapsed copy (quadric
collapses with reused vertices, 391,681 F) from 64 orthographic Fibonacci
views at 1024 x 1024 samples (screen 120 px, supersample 8), limit 1e3 so
every view is measured. Meshes are resident in HBM before timing starts.

Prints one JSON line with the throughput, the dominant kernel's roofline
(HIP-event kernel times from the library's own stream; algorithmic bytes:
geometry once, one bit per sample per mask, 8 B per sample of distance
transform) and a CPU baseline timed on a bounded sample of the
same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
COPY_PROBE = ROOT / "bench_calib" / "libblitz_copy_probe.so"
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (and matrix) peak
BF16_MFMA_PEAK_TFLOPS = 2500.0  # dense
FP32_PEAK_TFLOPS = 157.3
METRIC = "audited collapses/s & ms per LOD chain at ≤1px error; raster audit % HBM roof"


def workload():
    """The raster-audit configuration; the candidate is the reference's
    classical reduction (endpoint output) of the heightfield to 75 %, made by
    the device reducer (same triangles as the reference's reduce())."""
    from blitzremesher_amd import EvalSettings, meshes
    from blitzremesher_amd.audit import bounds, reduce_reuse_device
    ref = meshes.heightfield(512, 0.1, 2)
    cand = reduce_reuse_device(ref, int(round(ref.triangles * 0.75)))
    center, radius = bounds(ref)
    # The heightfield is an open surface seen from all around the sphere, so
    # both sides are drawn (force_two_sided); with one-sided faces the views
    # from below are empty or near-empty and the audit stops there.
    settings = EvalSettings(profile=0, orthographic=64, perspective=0, rotation_seed=0xB1172026, screen_size=120,
                            supersample=8, max_supersample=8, limit=1e3, max_changed_area=1.0,
                            force_two_sided=True)
    return ref, cand, center, radius, settings


def algorithmic_bytes(ref, cand, views: int, samples: int, distance_views: int) -> dict:
    """Bytes an ideal audit must move: geometry read once (12 B per vertex and per face of each mesh), one bit
    per sample written for each of the two masks per view, and for each view
    needing the distance transform two directions of 16-bit row distances
    written and read (8 B per sample)."""
    geometry = 12 * (ref.positions.shape[0] + ref.triangles) + 12 * (cand.positions.shape[0] + cand.triangles)
    masks = views * 2 * samples // 8
    # One projection per distinct vertex stream and view: 16 B read, 16 B written per vertex.
    shared = ref.positions.shape == cand.positions.shape and bool((ref.positions == cand.positions).all())
    project = 32 * views * (ref.positions.shape[0] + (0 if shared else cand.positions.shape[0]))
    return {"geometry": geometry, "masks": masks, "raster": geometry + masks, "project": project,
            "distance": distance_views * 8 * samples, "audit": geometry + 2 * masks + distance_views * 8 * samples}


def raster_fp64_ops(meshes_, center, radius, settings) -> int:
    """FP64 operations the reference's coverage raster performs for one audit
    (src/evaluate.cpp:144-170), both meshes over every view: per drawn face 19
    (signed area 7, three conservative radii 4 each) and per sample of its
    clipped bounding box 26 (sample centre 2, three edge functions 7 each,
    three radius additions); comparisons are not counted."""
    import numpy as np
    from blitzremesher_amd.audit import cameras
    cams = cameras(center, radius, settings.screen_size, settings.orthographic, settings.perspective,
                   settings.rotation_seed)
    size = int(-(-(settings.screen_size + 8) // 1)) * settings.supersample
    total = 0
    for mesh in meshes_:
        p = mesh.positions.astype(np.float64) - np.asarray(center, np.float64)
        idx = mesh.indices.reshape(-1, 3).astype(np.int64)
        for cam in cams:
            r, u = np.asarray(cam.right, np.float64), np.asarray(cam.up, np.float64)
            x = (p[:, 0] * r[0] + p[:, 1] * r[1] + p[:, 2] * r[2]) * cam.scale * settings.supersample + size * .5
            y = (p[:, 0] * u[0] + p[:, 1] * u[1] + p[:, 2] * u[2]) * cam.scale * settings.supersample + size * .5
            fx, fy = x[idx], y[idx]
            area = (fx[:, 1] - fx[:, 0]) * (fy[:, 2] - fy[:, 0]) - (fy[:, 1] - fy[:, 0]) * (fx[:, 2] - fx[:, 0])
            drawn = np.abs(area) >= 1e-16
            if not settings.force_two_sided:
                drawn &= area > 0
            x0 = np.clip(np.floor(fx.min(1)), 0, size - 1)
            x1 = np.clip(np.floor(fx.max(1)), 0, size - 1)
            y0 = np.clip(np.floor(fy.min(1)), 0, size - 1)
            y1 = np.clip(np.floor(fy.max(1)), 0, size - 1)
            box = (x1 - x0 + 1) * (y1 - y0 + 1)
            total += int(19 * drawn.sum() + 26 * box[drawn].sum())
    return total


_cpu_worker = {}


def _cpu_worker_init(ref, cand, center, radius, sample):
    _cpu_worker.update(ref=ref, cand=cand, center=center, radius=radius, sample=sample)


def _cpu_worker_audit(_=None) -> float:
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import Oracle
    w = _cpu_worker
    t0 = time.perf_counter()
    Oracle("ref").evaluate(w["ref"], w["cand"], w["center"], w["radius"], w["sample"])
    return time.perf_counter() - t0


def cpu_baseline_all_cores(ref, cand, center, radius, settings, sample_views: int) -> dict | None:
    """The reference CPU audit on every core at once: one worker process per
    core, each auditing the same `sample_views`-view sample; audits/s of the
    64-view workload = cores x sample / 64 / wall."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import available
    from blitzremesher_amd import EvalSettings
    if not available("ref"):
        return None
    import multiprocessing as mp
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    sample = EvalSettings(**{**settings.__dict__, "orthographic": sample_views})
    with mp.get_context("spawn").Pool(cores, initializer=_cpu_worker_init,
                                      initargs=(ref, cand, center, radius, sample)) as pool:
        pool.map(_cpu_worker_audit, range(cores))  # start-up and first touch
        t0 = time.perf_counter()
        pool.map(_cpu_worker_audit, range(cores))
        wall = time.perf_counter() - t0
    return {"value": cores * sample_views / settings.orthographic / wall, "unit": "audits/s", "cores": cores,
            "kind": "reference", "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "sample": f"{cores} processes, each a {sample_views}-view sample of the same audit, in {wall:.2f} s"}


def cpu_baseline(ref, cand, center, radius, settings, sample_views: int):
    """Times the reference CPU audit (oracle/_ref build) or, without it, the
    oracle restatement, on `sample_views` views of the same raster size, one
    thread; scales to audits/s of the 64-view workload."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import Oracle, available
    from blitzremesher_amd import EvalSettings
    kind = "reference" if available("ref") else "port"
    oracle = Oracle("ref" if kind == "reference" else "oracle")
    sample = EvalSettings(**{**settings.__dict__, "orthographic": sample_views})
    t0 = time.perf_counter()
    m = oracle.evaluate(ref, cand, center, radius, sample)
    dt = time.perf_counter() - t0
    per_view = dt / sample_views
    value = 1.0 / (per_view * settings.orthographic)
    return {"value": value, "unit": "audits/s", "cores": 1, "kind": kind,
            "sample": f"{sample_views} Fibonacci views (same 1024^2 raster, same meshes) in {dt:.2f} s, "
                      f"scaled to 64 views"}, sample, m


def scorer_leg(steps: int) -> dict:
    """Scorer forward on the 500 x 200 noisy torus (100,000 V / 600,000
    directed actions), FP32 compensated and BF16 MFMA, and one training step
    on an 8,192-action batch; kernel time from the library's HIP events."""
    import numpy as np
    from blitzremesher_amd import meshes
    from blitzremesher_amd.audit import bounds
    from blitzremesher_amd.scorer import BF16, FP32, Scorer, TrainSettings, he_weights
    mesh = meshes.torus(1.0, 0.3, 500, 200, 0.01, 3)
    c, r = bounds(mesh)
    out = {"workload": "noisy torus R=1 r=0.3 500x200 (100,000 V / 200,000 F / 600,000 directed actions), "
                       "32 features, 3 MP layers x 128, edge MLP 256-128-4, He-normal seed 4"}
    with Scorer(he_weights(4)) as s:
        g = s.graph(mesh, c, r)
        nv, na = g.vertices, g.actions
        # Direct formulation (per-action 256 -> 128 product) and the one the
        # BF16 path implements (edge layer factorised over the vertices).
        mp = 2 * nv * (64 * 128 + 256 * 128 + 256 * 128)
        flops = mp + 2 * na * (256 * 128 + 128 * 4)
        factorised = mp + 2 * nv * 128 * 256 + na * (2 * 128 + 2 * 128 * 4)
        out["flops_per_forward"] = flops
        out["flops_per_forward_factorised"] = factorised
        for name, p in (("fp32", FP32), ("bf16", BF16)):
            for _ in range(2):
                s.forward(g, p, download=False)
            ms = []
            for _ in range(steps):
                s.forward(g, p, download=False)
                ms.append(s.last_forward_ms())
            t = float(np.median(ms))
            peak = BF16_MFMA_PEAK_TFLOPS if p == BF16 else FP32_PEAK_TFLOPS
            done = factorised if p == BF16 else flops
            out[name] = {"ms": round(t, 4), "actions_per_s": round(na / (t * 1e-3), 1),
                         "formulation": "factorised edge layer" if p == BF16 else "direct",
                         "tflops": round(done / (t * 1e-3) / 1e12, 2), "peak_tflops": peak,
                         "frac": round(done / (t * 1e-3) / 1e12 / peak, 4),
                         "direct_equivalent_tflops": round(flops / (t * 1e-3) / 1e12, 2)}
        rng = np.random.default_rng(5)
        batch = 8192
        ids = rng.choice(na, batch, replace=False).astype(np.uint32)
        valid = rng.random(batch) < 0.85
        pref = valid & (rng.random(batch) < 0.25)
        labels = (valid * 1 | pref * 2 | (rng.random(batch) < 0.5) * 4).astype(np.uint8)
        targets = rng.uniform(-1, 1, (batch, 3)).astype(np.float32)
        settings = TrainSettings(pool=32, precision=BF16)
        for _ in range(2):
            s.train_step(g, ids, labels, targets, settings)
        ms = []
        for _ in range(steps):
            s.train_step(g, ids, labels, targets, settings)
            ms.append(s.last_step_ms())
        t = float(np.median(ms))
        out["error_vs_fp64"] = scorer_error()
        out["train_bf16"] = {"batch": batch, "ms_per_step": round(t, 4), "edges_per_s": round(batch / (t * 1e-3), 1),
                             "note": "full-graph forward (all 100,000 vertices), batch edge layer, FP32 backward, "
                                     "AdamW"}
    return out


def patch_leg(ref, center, radius, settings, count: int, repeats: int) -> dict:
    """The raster-audit workload's local patches: `count` seeded BFS regions
    of 16-200 faces of the heightfield, each re-meshed by legal interior
    collapses, audited against the bound heightfield over the same 64 views
    in one batch (dirty tiles re-rasterized, exact EDT)."""
    import numpy as np
    from blitzremesher_amd import Audit
    from blitzremesher_amd.audit import local_patches
    t0 = time.perf_counter()
    patches = local_patches(ref, count, 0xB1172026, 16, 200)
    generate_s = time.perf_counter() - t0
    removed = int(patches.removed_offsets[-1])
    added = int(patches.added_offsets[-1])
    with Audit(0) as audit:
        t0 = time.perf_counter()
        binding = audit.bind(ref, center, radius, settings)
        bind_ms = 1000 * (time.perf_counter() - t0)
        audit.evaluate_patches(binding, patches, settings)  # warm-up
        times, kernel = [], []
        for _ in range(repeats):
            t0 = time.perf_counter()
            ms, fallbacks = audit.evaluate_patches(binding, patches, settings)
            times.append(time.perf_counter() - t0)
            kernel.append(audit.last_timing()["patch_ms"])
        t = float(np.median(times))
    # An interior collapse removes the two faces of its edge.
    collapses = (removed - added) // 2
    return {"workload": f"{count} local patches (BFS regions of 16-200 faces, re-meshed by legal interior "
                        "collapses) of the 512x512 heightfield, 64 views at 1024^2, one batch",
            "patches": count, "removed_faces": removed, "added_faces": added, "collapses": collapses,
            "bind_ms": round(bind_ms, 2), "batch_ms": round(1000 * t, 3),
            "patch_kernel_ms": round(float(np.median(kernel)), 3),
            "patch_audits_per_s": round(count / t, 1),
            # One audit against the source per patch: not the four-gate count.
            "single_audit_collapses_per_s": round(collapses / t, 1),
            "full_path_fallbacks": fallbacks, "passed": int(sum(m.passed for m in ms)),
            "generate_s": round(generate_s, 2)}


def appearance_leg(steps: int) -> dict:
    """Normals and attributes profiles on the CPU-runnable icosphere with
    normal and colour streams (10,242 V / 20,480 F) against its 15,000-face
    quadric copy carrying the same streams: 16 views at 256 px, ss 8 (refined
    to 16 where uncertain), limits 1 px and 3 px; and the top-left overlap
    measure of the icosphere over the same views."""
    import numpy as np
    from blitzremesher_amd import Audit, EvalSettings, Mesh, meshes
    from blitzremesher_amd.audit import bounds, reduce_reuse
    ico = meshes.with_attributes(meshes.icosphere(5, 0.02, 1), 7)
    red = reduce_reuse(ico, 15000)
    cand = Mesh(red.positions, red.indices, normals=ico.normals, colors=ico.colors)
    center, radius = bounds(ico)
    out = {"workload": "icosphere 10,242 V / 20,480 F with normals and colours vs its 15,000-face quadric copy, "
                       "16 views at 256 px, ss 8 -> 16"}
    with Audit(0) as audit:
        dr, dc = audit.upload(ico), audit.upload(cand)
        for profile in (1, 2):
            for limit in (1.0, 3.0):
                s = EvalSettings(profile=profile, screen_size=256, supersample=8, max_supersample=16,
                                 orthographic=16, perspective=0, limit=limit)
                audit.evaluate(dr, dc, center, radius, s)
                times, app = [], []
                for _ in range(steps):
                    t0 = time.perf_counter()
                    m = audit.evaluate(dr, dc, center, radius, s)
                    times.append(time.perf_counter() - t0)
                    app.append(audit.last_timing()["appearance_ms"])
                out[f"profile{profile}_limit{limit:g}"] = {
                    "ms": round(1000 * float(np.median(times)), 3), "appearance_ms": round(float(np.median(app)), 3),
                    "error": m.error, "normal_degrees": round(m.normal_degrees, 4), "views": m.views_evaluated,
                    "passed": m.passed}
        audit.overlap(dr, center, radius, 256, 16, 0)
        t0 = time.perf_counter()
        value = audit.overlap(dr, center, radius, 256, 16, 0)
        out["overlap"] = {"value": value, "ms": round(1000 * (time.perf_counter() - t0), 3)}
    return out


def executor_leg(budget: int, cpu_budget: int) -> dict:
    """The plain classical method's audited collapses on the CPU-runnable
    icosphere (10,242 V / 20,480 F, radial noise 0.02): execute_actions with
    the current-plane order, batches of 64 independent collapses, each trial
    gated by an exact coverage audit against the source from 16 Fibonacci
    views at 256 px (ss 8, refinement to 16, limit 1 px), towards half the
    triangles. The reference's own CPU loop (oracle/_ref build, one thread)
    runs the first `cpu_budget` trials of the same run; the GPU repeats those
    trials for a parity check of output and work counters."""
    from blitzremesher_amd import Audit, EvalSettings, meshes
    from blitzremesher_amd.actions import execute_actions
    from blitzremesher_amd.audit import bounds
    mesh = meshes.icosphere(5, 0.02, 1)
    center, radius = bounds(mesh)
    gate = EvalSettings(profile=0, screen_size=256, supersample=8, max_supersample=16, orthographic=16,
                        perspective=0, rotation_seed=0xB1172026, limit=1.0)
    target = mesh.triangles // 2
    # Every gate audits against the same uploaded source: its masks are kept
    # across trials (cache_rasters), as the reference's RasterMemo keeps them.
    with Audit(0, cache_rasters=True) as audit:
        execute_actions(audit, mesh, center, radius, gate, target, 2, 4, 64)  # warm-up
        out, st = execute_actions(audit, mesh, center, radius, gate, target, budget, 4, 64)
        sample_out, sample_st = execute_actions(audit, mesh, center, radius, gate, target, cpu_budget, 4, 64)
    seconds = st["ms"] / 1000.0
    leg = {"workload": "execute_actions (current-plane order, batch 64) on the 10,242-vertex icosphere towards "
                       f"{target} triangles, each trial gated by a coverage audit (16 views, S=256, ss 8->16, "
                       "limit 1 px) against the source",
           "budget": budget, "trials": st["trials"], "accepted": st["accepted"], "rejected": st["rejected"],
           "ranked": st["ranked"], "states": st["states"], "triangles": int(len(out)), "ms": round(st["ms"], 2),
           "ms_per_trial": round(st["ms"] / max(1, st["trials"]), 3),
           "stage_ms": {k: round(st[f"{k}_ms"], 2) for k in ("build", "rank", "select", "trial", "gate")},
           # One gate (an audit against the source) per trial: not the four-gate count of the chain leg.
           "one_gate_collapses_per_s": round(st["accepted"] / seconds, 1) if seconds else None}
    try:
        sys.path.insert(0, str(ROOT / "tests"))
        from oracle_lib import available, ref_execute_actions
        if cpu_budget and available("ref"):
            t0 = time.perf_counter()
            cpu_out, cpu_st = ref_execute_actions(mesh, center, radius, gate, target, cpu_budget, 4, 64)
            dt = time.perf_counter() - t0
            same = (len(cpu_out) == len(sample_out) and (cpu_out == sample_out).all() and
                    all(cpu_st[k] == sample_st[k] for k in ("ranked", "trials", "accepted", "rejected")))
            leg["cpu_reference"] = {"trials": cpu_st["trials"], "accepted": cpu_st["accepted"], "s": round(dt, 2),
                                    "one_gate_collapses_per_s": round(cpu_st["accepted"] / dt, 1), "cores": 1,
                                    "gpu_same_trials_ms": round(sample_st["ms"], 2),
                                    "identical_output_and_counters": bool(same)}
    except Exception as e:  # the CPU comparison is optional
        leg["cpu_reference"] = {"error": str(e)}
    return leg


def neural_leg(train_steps: int, budget: int) -> dict:
    """The learned proposer end to end on the CPU-runnable icosphere: every
    legal collapse of the source is audited as a patch of the bound source
    (16 views at 256 px, ss 8, limit 1 px: accept/reject labels); the
    edge-graph scorer (He-normal init, seed 4) takes `train_steps` BF16 AdamW
    steps of the pairwise ranking loss on pools of 32 labelled collapses
    (batch 8,192); then execute_actions runs with the learned order, the
    untrained scorer's order and the plain current-plane order, `budget` trials
    each, every trial gated by the executor leg's exact audit."""
    import numpy as np
    from blitzremesher_amd import Audit, EvalSettings, meshes
    from blitzremesher_amd.actions import ActionSet, execute_actions
    from blitzremesher_amd.audit import bounds
    from blitzremesher_amd.scorer import BF16, PREFERRED, VALID, Scorer, TrainSettings, he_weights
    mesh = meshes.icosphere(5, 0.02, 1)
    center, radius = bounds(mesh)
    gate = EvalSettings(profile=0, screen_size=256, supersample=8, max_supersample=16, orthographic=16,
                        perspective=0, rotation_seed=0xB1172026, limit=1.0)
    label_gate = EvalSettings(**{**gate.__dict__, "max_supersample": 8})
    target = mesh.triangles // 2
    out = {"workload": "labels: every legal collapse of the 10,242-vertex icosphere audited as a patch (16 views, "
                       "S=256, ss 8, limit 1 px); training: BF16 ranking loss, batch 8,192 in pools of 32; executor: "
                       f"{budget} trials towards {target} triangles, batch 64, gate as the executor leg"}
    with Audit(0, cache_rasters=True) as audit, Scorer(he_weights(4)) as scorer, \
            Scorer(he_weights(4)) as untrained, ActionSet(mesh) as acts:
        n = len(acts.legal)
        t0 = time.perf_counter()
        patches = acts.patches(np.arange(n, dtype=np.uint32))
        binding = audit.bind(mesh, center, radius, label_gate)
        ms, fallbacks = audit.evaluate_patches(binding, patches, label_gate)
        label_s = time.perf_counter() - t0
        passed = np.array([m.passed for m in ms], bool)
        g = scorer.graph(mesh, center, radius)
        pairs = g.actions_array()
        keys = pairs[:, 0].astype(np.uint64) << np.uint64(32) | pairs[:, 1].astype(np.uint64)
        legal_keys = acts.legal[:, 0].astype(np.uint64) << np.uint64(32) | acts.legal[:, 1].astype(np.uint64)
        idx = np.searchsorted(keys, legal_keys).astype(np.uint32)
        assert np.array_equal(keys[idx], legal_keys)
        rng = np.random.default_rng(7)
        ts = TrainSettings(precision=BF16)
        losses = []
        t0 = time.perf_counter()
        for _ in range(train_steps):
            sel = rng.choice(n, 8192, replace=False)
            labels = (VALID | np.where(passed[sel], PREFERRED, 0)).astype(np.uint8)
            losses.append(scorer.train_step(g, idx[sel], labels, settings=ts).loss)
        train_s = time.perf_counter() - t0
        logits = scorer.forward(g, BF16)[:, 0][idx]
        top = np.argsort(-logits, kind="stable")[: n // 10]
        out["labels"] = {"legal": n, "passed": int(passed.sum()), "patch_audit_s": round(label_s, 3),
                         "full_path_fallbacks": int(fallbacks),
                         "labels_per_s": round(n / label_s, 1)}
        out["training"] = {"steps": train_steps, "s": round(train_s, 3), "loss_first": round(losses[0], 4),
                           "loss_last": round(float(np.mean(losses[-10:])), 4),
                           "pass_rate_top_decile": round(float(passed[top].mean()), 4),
                           "pass_rate_all": round(float(passed.mean()), 4)}
        g.release()
        for name, ranking, sc in (("current_plane", 4, None), ("learned", 0, scorer), ("untrained", 0, untrained)):
            o, st = execute_actions(audit, mesh, center, radius, gate, target, budget, ranking, 64, scorer=sc)
            out[name] = {"trials": st["trials"], "accepted": st["accepted"], "rejected": st["rejected"],
                         "triangles": int(len(o)), "ms": round(st["ms"], 1),
                         "accepted_per_trial": round(st["accepted"] / max(1, st["trials"]), 3),
                         "one_gate_collapses_per_s": round(st["accepted"] / (st["ms"] / 1000.0), 1)}
    return out


def training_leg(steps: int, count: int, pools_per_mesh: int, holdout: int = 8, chain_repeats: int = 2,
                 batched_meshes: int = 8, batched_workers: int = 4) -> dict:
    """Scorer training from method-generated labels: `count` seeded meshes of
    uneven size (2k-100k triangles, log-uniform; spheres, tori, extruded
    boxes, heightfields, seed 5); up to `pools_per_mesh` teacher pools of 16
    collapses per mesh, half from the classical current-plane order and half
    random, each collapse audited as a patch under the first chain level's
    four gates (the chain leg's settings: 16 views at 256 px, ss 4 search /
    8 audit, 1 px) and labelled with the reference's preferred sets; placement
    targets from the teacher solve. All graphs are joined into one device
    graph; each BF16 step (AdamW, lr 1e-3, clip 1) takes 8,192 rows = 512
    whole pools drawn from the first count - holdout meshes, so a batch spans
    many meshes. Reported: label time, the loss curve, step time, and the
    share of pools whose top-scored valid collapse is preferred, before and
    after training, on trained and held-out meshes."""
    import numpy as np
    from blitzremesher_amd import Audit, meshes
    from blitzremesher_amd.scorer import BF16, Scorer, TrainSettings, he_weights
    from blitzremesher_amd.teacher import build_dataset, sample_batch, top1_preferred
    settings = c1_chain_settings()
    t0 = time.perf_counter()
    family_meshes = meshes.training_set(count, 2000, 100000, 5)
    gen_s = time.perf_counter() - t0
    ms = [m for _, m in family_meshes]
    tris = np.array([m.triangles for m in ms])
    out = {"workload": f"{count} procedural meshes (log-uniform 2k-100k triangles, seed 5; sphere, torus, extruded "
                       f"box, heightfield), up to {pools_per_mesh} teacher pools of 16 per mesh audited under the "
                       "level-1 four gates (16 views, S=256, ss 4/8, 1 px); BF16 training, batch 8,192 rows "
                       f"(512 pools) spanning meshes, {steps} AdamW steps; last {holdout} meshes held out",
           "meshes": {"count": count, "triangles_total": int(tris.sum()), "triangles_min": int(tris.min()),
                      "triangles_max": int(tris.max()), "generate_s": round(gen_s, 3)}}
    with Audit(0) as audit, Scorer(he_weights(4)) as scorer:
        t0 = time.perf_counter()
        data = build_dataset(audit, scorer, ms, settings, pools_per_mesh)
        label_s = time.perf_counter() - t0
        rows = len(data.actions)
        valid = (data.bits & 1) != 0
        out["labels"] = {"pools": data.pools, "rows": rows, "valid": int(valid.sum()),
                         "preferred": int(((data.bits & 2) != 0).sum()), "placed": int(((data.bits & 4) != 0).sum()),
                         "full_path_fallbacks": int(data.fallbacks), "s": round(label_s, 3),
                         "row_audits_per_s": round(4 * rows / label_s, 1),
                         "joined_graph": {"vertices": data.graph.vertices, "actions": data.graph.actions}}
        train_pools = np.flatnonzero(data.pool_mesh < count - holdout)
        held_pools = np.flatnonzero(data.pool_mesh >= count - holdout)
        before = scorer.forward(data.graph, BF16)[:, 0]
        ts = TrainSettings(pool=16, precision=BF16)
        rng = np.random.default_rng(11)
        losses, step_ms = [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            a, b, t = sample_batch(data, train_pools, 8192, rng)
            r = scorer.train_step(data.graph, a, b, t, ts)
            losses.append(r.loss)
            step_ms.append(scorer.last_step_ms())
        train_s = time.perf_counter() - t0
        after = scorer.forward(data.graph, BF16)[:, 0]
        nv, nb = data.graph.vertices, 8192
        # Forward flops of the formulation run (every vertex through the three
        # message-passing layers, the batch through the edge MLP), backward
        # counted as twice the forward.
        flops = 3 * (2 * nv * (64 + 256 + 256) * 128 + 2 * nb * (256 * 128 + 128 * 4))
        med = float(np.median(step_ms))
        every = max(1, steps // 10)
        out["training"] = {
            "steps": steps, "s": round(train_s, 3), "step_ms_median": round(med, 3),
            "rows_per_s": round(nb / (med * 1e-3), 1),
            "flops_per_step": flops, "tflops": round(flops / (med * 1e-3) / 1e12, 2),
            "loss_curve": [[i, round(float(np.mean(losses[i:i + every])), 4)] for i in range(0, steps, every)],
            "top1_preferred": {"trained_before": round(top1_preferred(data, before, train_pools), 4),
                               "trained_after": round(top1_preferred(data, after, train_pools), 4),
                               "held_out_before": round(top1_preferred(data, before, held_pools), 4),
                               "held_out_after": round(top1_preferred(data, after, held_pools), 4)}}
        data.graph.release()
        out["chain_learned_vs_plain"] = learned_chain(scorer, chain_repeats)
        if batched_meshes:
            out["batched_chains"] = batched_chains(scorer.weights(), batched_meshes, batched_workers)
    return out


def c5_chain_settings():
    """The batched-generation chain: 6 levels from 512 px to 16 px,
    transition and LOD0 bounds 1 px, coverage profile, 32 orthographic search
    and audit views, ss 4 search / 8 audit / 16 max, candidate budget 8, beam
    2, max changed area 1."""
    from blitzremesher_amd.chain import ChainSettings
    return ChainSettings(levels=6, profile=0, base_pixels=512, last_pixels=16, max_lod0_delta_px=1.0,
                         transition=((0, 1), (1, 1)), search_orthographic=32, search_perspective=0,
                         search_seed=0xB1172024, audit_orthographic=32, audit_perspective=0, audit_seed=0xB1172026,
                         search_supersample=4, audit_supersample=8, max_supersample=16, candidate_budget=8,
                         beam_width=2, max_changed_area=1.0)


def batched_chains(weights, count: int, workers: int) -> dict:
    """Batched end-to-end generation on a bounded sample of the 128-mesh set
    (1k-250k triangles, log-uniform, seed 11; spheres, tori, extruded boxes,
    heightfields): `count` meshes, every 8th by size from the smallest, each
    a 6-level chain (c5_chain_settings) with `workers` chains in flight (one
    audit object, and for learned ranking one scorer with the trained
    weights, per worker thread). Learned (BF16 logits) and the plain
    current-plane order run the same meshes; reported per ranking: wall time,
    four-gate audited collapses (and per second), trials and the emitted
    triangles summed over the meshes per level."""
    import threading
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    from blitzremesher_amd import Audit, meshes
    from blitzremesher_amd.chain import ChainOptions, generate
    from blitzremesher_amd.scorer import Scorer
    everything = meshes.training_set(128, 1000, 250000, seed=11)
    order = np.argsort([m.triangles for _, m in everything], kind="stable")
    pick = [int(i) for i in order[0:8 * count:8]]
    settings = c5_chain_settings()
    sizes = [everything[i][1].triangles for i in pick]
    out = {"workload": f"{len(pick)} of the 128 procedural meshes (every 8th by size from the smallest: "
                       f"{min(sizes)}-{max(sizes)} triangles; set of 1k-250k, seed 11), 6-level chains "
                       "(512 -> 16 px, 1 px bounds, 32 views, ss 4/8/16, budget 8, beam 2; 64 trials x batch 32 "
                       f"per proposal), {workers} chains in flight",
           "meshes": [{"family": everything[i][0], "triangles": everything[i][1].triangles} for i in pick]}
    # The raster tuner's first launches of each shape, outside the timing.
    with Audit(0, cache_rasters=True) as audit:
        generate(audit, everything[pick[0]][1], settings,
                 ChainOptions(action_trials=64, action_batch=32, ranking=4, overdraw_tiebreak=True))
    for name, ranking in (("learned", 0), ("plain", 4)):
        local = threading.local()
        owned = []
        lock = threading.Lock()

        def run(i):
            if not hasattr(local, "audit"):
                local.audit = Audit(0, cache_rasters=True)
                local.scorer = Scorer(weights) if ranking == 0 else None
                with lock:
                    owned.append((local.audit, local.scorer))
            options = ChainOptions(action_trials=64, action_batch=32, ranking=ranking, overdraw_tiebreak=True)
            return generate(local.audit, everything[i][1], settings, options, scorer=local.scorer)

        t0 = time.perf_counter()
        with ThreadPoolExecutor(workers) as pool:
            results = list(pool.map(run, pick))
        wall = time.perf_counter() - t0
        for audit, scorer in owned:
            audit.close()
            if scorer is not None:
                scorer.close()
        collapses = sum(r.info["legal_collapses"] for r in results)
        out[name] = {"wall_s": round(wall, 2), "four_gate_audited_collapses": collapses,
                     "four_gate_audited_collapses_per_s": round(collapses / wall, 1),
                     "trials": sum(r.info["action_trials"] for r in results),
                     "rejected_collapses": sum(r.info["rejected_collapses"] for r in results),
                     "triangles_per_level": [sum(r.lods[k].triangles for r in results) for k in range(6)],
                     "stage_ms_summed": {k: round(sum(r.info[f"{k}_ms"] for r in results), 1)
                                         for k in ("executor", "gate", "select", "trial", "speculative",
                                                   "speculative_build", "offer", "confirm")},
                     "levels_passed": all(l.source.passed and l.adjacent.passed for r in results for l in r.lods[1:]),
                     "status": sorted({int(r.status) for r in results})}
    lp, pp = out["learned"], out["plain"]
    out["ratio_collapses_per_s"] = round(lp["four_gate_audited_collapses_per_s"] /
                                         max(pp["four_gate_audited_collapses_per_s"], 1e-9), 3)
    out["learned_triangles_at_most_plain"] = all(a <= b for a, b in zip(lp["triangles_per_level"],
                                                                        pp["triangles_per_level"]))
    return out


def learned_chain(scorer, repeats: int) -> dict:
    """The chain leg's 4-level chain of the 10,242-vertex icosphere (not one
    of the training meshes) with the executor's collapses ranked by the
    trained scorer's BF16 logits (ranking 0) and by the plain current-plane
    order (ranking 4), same gates and budgets: ms per chain, four-gate
    audited collapses (and per second), trials, rejected collapses and the
    triangles of every level."""
    import numpy as np
    from blitzremesher_amd import Audit, meshes
    from blitzremesher_amd.chain import ChainOptions, generate
    mesh = meshes.icosphere(5, 0.02, 1)
    settings = c1_chain_settings()
    out = {}
    for name, ranking in (("learned", 0), ("plain", 4)):
        options = ChainOptions(action_trials=64, action_batch=32, ranking=ranking, overdraw_tiebreak=True)
        with Audit(0, cache_rasters=True) as audit:
            generate(audit, mesh, settings, options, scorer=scorer)  # warm-up
            runs = [generate(audit, mesh, settings, options, scorer=scorer) for _ in range(repeats)]
        r = runs[-1]
        ms = float(np.median([x.info["ms"] for x in runs]))
        i = r.info
        out[name] = {"ms_per_chain": round(ms, 1), "triangles": [l.triangles for l in r.lods],
                     "four_gate_audited_collapses": i["legal_collapses"],
                     "four_gate_audited_collapses_per_s": round(i["legal_collapses"] / (ms * 1e-3), 1),
                     "rejected_collapses": i["rejected_collapses"], "trials": i["action_trials"],
                     "stage_ms": {k: round(i[f"{k}_ms"], 1) for k in ("executor", "gate", "rank", "speculative")},
                     "levels_passed": all(l.source.passed and l.adjacent.passed for l in r.lods[1:]),
                     "identical_runs": all([l.triangles for l in x.lods] == [l.triangles for l in r.lods]
                                           for x in runs)}
    lp, pp = out["learned"], out["plain"]
    out["ratio_collapses_per_s"] = round(lp["four_gate_audited_collapses_per_s"] /
                                         max(pp["four_gate_audited_collapses_per_s"], 1e-9), 3)
    out["learned_triangles_at_most_plain"] = all(a <= b for a, b in zip(lp["triangles"], pp["triangles"]))
    out["target_ratio"] = 10.0
    return out


def c1_chain_settings():
    """The CPU-runnable chain configuration: 4 levels from 256 px to 32 px,
    transition and LOD0 bounds 1 px, coverage profile, 16 orthographic search
    and audit views (reference default seeds), ss 4 search / 8 audit / 16 max,
    candidate budget 8, beam 2, max changed area 1."""
    from blitzremesher_amd.chain import ChainSettings
    return ChainSettings(levels=4, profile=0, base_pixels=256, last_pixels=32, max_lod0_delta_px=1.0,
                         transition=((0, 1), (1, 1)), search_orthographic=16, search_perspective=0,
                         search_seed=0xB1172024, audit_orthographic=16, audit_perspective=0, audit_seed=0xB1172026,
                         search_supersample=4, audit_supersample=8, max_supersample=16, candidate_budget=8,
                         beam_width=2, max_changed_area=1.0)


def _cpu_classical_chain(_=None) -> float:
    """One reference classical chain (oracle/_ref generate) of the icosphere;
    returns its wall seconds."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import ChainSettings as RefChain
    from oracle_lib import Oracle
    from blitzremesher_amd import meshes
    s = c1_chain_settings()
    cs = RefChain(levels=s.levels, beam_width=s.beam_width, search_supersample=s.search_supersample,
                  audit_supersample=s.audit_supersample, max_supersample=s.max_supersample,
                  candidate_budget=s.candidate_budget, search_ortho=s.search_orthographic, search_persp=0,
                  audit_ortho=s.audit_orthographic, audit_persp=0, search_seed=s.search_seed, audit_seed=s.audit_seed,
                  base_pixels=s.base_pixels, last_pixels=s.last_pixels, max_lod0_delta_px=s.max_lod0_delta_px,
                  transition0=1.0, transition1=1.0, max_changed_area=s.max_changed_area)
    mesh = meshes.icosphere(5, 0.02, 1)
    t0 = time.perf_counter()
    Oracle("ref").generate(mesh, cs)
    return time.perf_counter() - t0


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def chain_cpu_baseline() -> dict | None:
    """The reference's classical CPU generator (QEM reduce proposals, CPU
    coverage audits; oracle/_ref generate) on the chain leg's mesh and
    settings: one chain on one core, then one chain per core on all cores at
    once (worker processes; run before the GPU is touched)."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import available
    if not available("ref"):
        return None
    import multiprocessing as mp
    one = _cpu_classical_chain()
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(cores) as pool:
        pool.map(_cpu_classical_chain, range(cores))
    wall = time.perf_counter() - t0
    return {"ms_per_chain_1_core": round(1000 * one, 1), "chains_per_s_all_cores": round(cores / wall, 3),
            "ms_per_chain_all_cores": round(1000 * wall / cores, 1), "cores": cores, "cpu_model": cpu_model(),
            "nproc": os.cpu_count(),
            "note": "reference generate(): QEM reduce proposals + CPU coverage audits, same mesh and settings"}


def chains_in_flight(mesh, settings, options, workers: int, chains: int, single) -> dict:
    """`chains` copies of one chain run `workers` at a time, one audit object
    per worker thread: chains per second of the whole set, and whether every
    chain equals the single-threaded one."""
    import threading
    from concurrent.futures import ThreadPoolExecutor
    from blitzremesher_amd import Audit
    from blitzremesher_amd.chain import generate
    local = threading.local()
    owned = []
    lock = threading.Lock()

    def run(_):
        if not hasattr(local, "audit"):
            local.audit = Audit(0, cache_rasters=True)
            with lock:
                owned.append(local.audit)
        return generate(local.audit, mesh, settings, options)

    with ThreadPoolExecutor(workers) as pool:  # one untimed round
        list(pool.map(run, range(workers)))
        t0 = time.perf_counter()
        results = list(pool.map(run, range(chains)))
        wall = time.perf_counter() - t0
    for audit in owned:
        audit.close()
    same = all([l.indices.tobytes() for l in x.lods] == [l.indices.tobytes() for l in single.lods] for x in results)
    return {"chains": chains, "wall_s": round(wall, 2), "chains_per_s": round(chains / wall, 3),
            "ms_per_chain_amortized": round(1000 * wall / chains, 1),
            "four_gate_audited_collapses_per_s": round(sum(x.info["legal_collapses"] for x in results) / wall, 1),
            "identical_to_single": same}


def chain_leg(repeats: int, cpu_reference: dict | None) -> dict:
    """LOD chains of the CPU-runnable icosphere (10,242 V / 20,480 F, radial
    noise 0.02, seed 1) under c1_chain_settings: explicit reuse output,
    proposals from the source by the collapse executor in the current-plane
    order (the plain classical order on the device; 64 trials per proposal,
    batches of 32), every trial certified by the four gates (source search,
    adjacent search, source audit, adjacent audit), every offer measured by
    the four gates and the selected chain confirmed by full audits, overdraw
    tie-break on. `four_gate_audited_collapses` counts collapses that passed
    all four gates."""
    import numpy as np
    from blitzremesher_amd import Audit, meshes
    from blitzremesher_amd.chain import ChainOptions, generate
    mesh = meshes.icosphere(5, 0.02, 1)
    settings = c1_chain_settings()
    options = ChainOptions(action_trials=64, action_batch=32, ranking=4, overdraw_tiebreak=True)
    with Audit(0, cache_rasters=True) as audit:
        generate(audit, mesh, settings, options)  # warm-up
        runs = [generate(audit, mesh, settings, options) for _ in range(repeats)]
    r = runs[-1]
    ms = float(np.median([x.info["ms"] for x in runs]))
    i = r.info
    in_flight = {str(w): chains_in_flight(mesh, settings, options, w, 2 * w, r) for w in (4, 8)}
    return {"workload": "4-level LOD chain (256 -> 32 px, transition and LOD0 bounds 1 px, 16 views, ss 4/8/16, "
                        "budget 8, beam 2) of the 10,242-vertex icosphere; device collapse executor in "
                        "current-plane order (64 trials x batch 32 per proposal), four-gate certified trials, "
                        "GPU audits",
            "ms_per_chain": round(ms, 1), "chains": repeats,
            "triangles": [l.triangles for l in r.lods],
            "four_gate_audited_collapses": i["legal_collapses"],
            "four_gate_audited_collapses_per_s": round(i["legal_collapses"] / (ms * 1e-3), 1),
            "rejected_collapses": i["rejected_collapses"], "trials": i["action_trials"], "proposals": i["decoded"],
            "gate_audits": i["gate_audits"], "offer_audits": int(sum(i["audit_evaluations"])),
            "confirm_audits": i["confirm_audits"], "finalists": len(r.finalists), "status": r.status,
            "stage_ms": {k: round(i[f"{k}_ms"], 1) for k in ("executor", "gate", "offer", "confirm", "tiebreak")},
            "executor_stage_ms": {k: round(i[f"{k}_ms"], 1) for k in ("build", "rank", "select", "trial",
                                                                      "speculative")},
            "levels_passed": all(l.source.passed and l.adjacent.passed for l in r.lods[1:]),
            "chains_in_flight": in_flight,
            "cpu_reference_classical": cpu_reference}


def reducer_leg(repeats: int) -> dict:
    """The classical reducer (the reference's reduce(), endpoint output) of
    the 512 x 512 heightfield to 75 % on the device, beside the reference
    build's own reduce() on one core (same triangles, checked)."""
    import numpy as np
    from blitzremesher_amd import meshes
    from blitzremesher_amd.audit import reduce_reuse_device
    mesh = meshes.heightfield(512, 0.1, 2)
    target = int(round(mesh.triangles * 0.75))
    reduce_reuse_device(mesh, target)  # warm-up
    times, report = [], {}
    for _ in range(repeats):
        t0 = time.perf_counter()
        got = reduce_reuse_device(mesh, target, report=report)
        times.append(time.perf_counter() - t0)
    ms = 1000 * float(np.median(times))
    out = {"workload": "reduce() with endpoint output of the 512x512 heightfield (522,242 F) to 75 %",
           "ms": round(ms, 2), "triangles": got.triangles, "collapses": report.get("collapses"),
           "passes": report.get("passes"), "greedy_rounds": report.get("rounds"),
           "collapses_per_s": round(report.get("collapses", 0) / (ms * 1e-3), 1)}
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import Oracle, available
    if available("ref"):
        t0 = time.perf_counter()
        want = Oracle("ref").reduce_reuse(mesh, target)
        cpu = time.perf_counter() - t0
        out["cpu_reference"] = {"ms": round(1000 * cpu, 1), "cores": 1, "kind": "reference",
                                "identical_triangles": bool(np.array_equal(want.indices, got.indices))}
    return out


def copy_probe(mib: int) -> int:
    """Runs the counter-calibration copy kernel (see profiles/) and prints its
    bandwidth."""
    import ctypes as C
    lib = C.CDLL(str(COPY_PROBE))
    ms = C.c_double()
    rc = lib.blitz_copy_probe(0, C.c_size_t(mib << 20), 5, C.byref(ms))
    if rc:
        print(f"copy probe failed ({rc})", file=sys.stderr)
        return 1
    print(json.dumps({"copy_bytes": mib << 20, "ms": ms.value,
                      "gbs": round(2 * (mib << 20) / (ms.value * 1e-3) / 1e9, 1)}), flush=True)
    return 0


def scorer_error() -> dict | None:
    """BF16 / FP32 forward error of the 100,000-vertex torus against the FP64
    oracle, from the committed summary of the GPU test that measures it."""
    path = ROOT / "profiles" / "scorer_torus100k_error.md"
    if not path.exists():
        return None
    out = {}
    for line in path.read_text().splitlines():
        for key in ("bf16_max_rel", "bf16_rms_rel", "fp32_max_rel"):
            if line.startswith(key + ":"):
                out[key] = float(line.split(":", 1)[1])
    out["source"] = "profiles/scorer_torus100k_error.md (tests/test_configs_gpu.py, all 600,000 actions)"
    return out


def measured_traffic() -> int | None:
    """HBM bytes per raster launch from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel calibration), if any."""
    path = ROOT / "profiles" / "raster_band_kernel_heightfield512_counters.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":", 1)[1].strip())
    return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-sample-views", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-scorer", action="store_true")
    ap.add_argument("--no-patches", action="store_true")
    ap.add_argument("--patches", type=int, default=4096)
    ap.add_argument("--no-executor", action="store_true")
    ap.add_argument("--no-appearance", action="store_true")
    ap.add_argument("--no-neural", action="store_true")
    ap.add_argument("--no-chain", action="store_true")
    ap.add_argument("--no-training", action="store_true")
    ap.add_argument("--no-reducer", action="store_true")
    ap.add_argument("--training-steps", type=int, default=200)
    ap.add_argument("--training-meshes", type=int, default=64)
    ap.add_argument("--training-pools", type=int, default=96)
    ap.add_argument("--chain-repeats", type=int, default=2)
    ap.add_argument("--batched-meshes", type=int, default=8)
    ap.add_argument("--batched-workers", type=int, default=4)
    ap.add_argument("--neural-train-steps", type=int, default=300)
    ap.add_argument("--executor-budget", type=int, default=400)
    ap.add_argument("--executor-cpu-budget", type=int, default=3)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run the counter-calibration copy of this many MiB")
    args = ap.parse_args()
    if args.copy_probe_mib:
        return copy_probe(args.copy_probe_mib)
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} (WORLD_SIZE {world}) is not supported: every box here "
              f"has one GPU and the audit path has no multi-GPU leg", file=sys.stderr)
        return 2

    from blitzremesher_amd import Audit, _lib
    import numpy as np
    started = time.perf_counter()

    def progress(step: str) -> None:
        print(f"bench: {step} done at {time.perf_counter() - started:.1f} s", file=sys.stderr, flush=True)

    # CPU-only baselines run first, before this process touches the GPU.
    chain_cpu = None if (args.no_cpu or args.no_chain) else chain_cpu_baseline()
    progress("CPU chain baseline")
    ref, cand, center, radius, settings = workload()
    progress("workload")
    # torch is not imported: its wheel bundles a second HIP runtime, and two
    # HIP runtimes in one process cannot share the device. The library's own
    # hipDeviceSynchronize brackets the timed region instead.
    lib = _lib.load()

    def sync() -> None:
        _lib.check(lib.blitz_hip_device_synchronize(0))

    with Audit(0) as audit:
        dref, dcand = audit.upload(ref), audit.upload(cand)
        m = None
        for _ in range(args.warmup):
            m = audit.evaluate(dref, dcand, center, radius, settings)
        sync()
        raster_ms = compare_ms = distance_ms = band_ms = project_ms = split_ms = 0.0
        launches = band_launches = 0
        distance_views = 0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            m = audit.evaluate(dref, dcand, center, radius, settings)
            t = audit.last_timing()
            raster_ms += t["raster_ms"]
            band_ms += t["band_ms"]
            project_ms += t["project_ms"]
            split_ms += t["split_ms"]
            band_launches += t["band_launches"]
            compare_ms += t["compare_ms"]
            distance_ms += t["distance_ms"]
            launches += t["passes"]
            distance_views = t["distance_views"]
        sync()
        elapsed = time.perf_counter() - t0
        ms_per_step = 1000.0 * elapsed / args.steps
        audit_chunks = audit.last_timing()["raster_chunks"]
        samples = (int(-(-(settings.screen_size + 8) // 1)) * settings.supersample) ** 2
        views = settings.orthographic + settings.perspective
        b = algorithmic_bytes(ref, cand, views, samples, views)
        # Device time per launch of each kernel of the audit (HIP events around
        # the launches on the library's stream); the raster pass also holds
        # the mask clears and the status read-back.
        kernels = {"raster_band_kernel": band_ms / band_launches, "project_kernel": project_ms / band_launches,
                   "compare_kernel": compare_ms / launches, "distance_kernels": distance_ms / launches}
        dominant = max(kernels, key=kernels.get)
        # The raster's secondary (compute) roof: the reference algorithm's FP64
        # operations per launch over the FP64 vector peak.
        fp64_ops = raster_fp64_ops((ref, cand), center, radius, settings)
        fp64_tflops = fp64_ops / (kernels["raster_band_kernel"] * 1e-3) / 1e12
        per_launch_bytes = {"raster_band_kernel": b["raster"], "project_kernel": b["project"],
                            "compare_kernel": 2 * b["masks"], "distance_kernels": b["distance"]}[dominant]
        achieved = per_launch_bytes / (kernels[dominant] * 1e-3) / 1e9
        gpu_sample = None
        cpu = None
        parity = None
        if not args.no_cpu:
            cpu, sample, cpu_m = cpu_baseline(ref, cand, center, radius, settings, args.cpu_sample_views)
            gpu_sample = audit.evaluate(dref, dcand, center, radius, sample)
            parity = gpu_sample.as_tuple() == cpu_m.as_tuple()
            cpu_all = cpu_baseline_all_cores(ref, cand, center, radius, settings, 8)
            if cpu_all:
                cpu["all_cores"] = cpu_all
    progress("raster audit")

    def leg(skip: bool, name: str, run):
        if skip:
            return None
        result = run()
        progress(name)
        return result

    reps = max(3, min(args.steps, 10))
    scorer = leg(args.no_scorer, "scorer", lambda: scorer_leg(reps))
    patches = leg(args.no_patches, "patches",
                  lambda: patch_leg(ref, center, radius, settings, args.patches, reps))
    appearance = leg(args.no_appearance, "appearance", lambda: appearance_leg(reps))
    neural = leg(args.no_neural, "neural", lambda: neural_leg(args.neural_train_steps, args.executor_budget))
    chain = leg(args.no_chain, "chain", lambda: chain_leg(args.chain_repeats, chain_cpu))
    training = leg(args.no_training, "training",
                   lambda: training_leg(args.training_steps, args.training_meshes, args.training_pools,
                                        chain_repeats=args.chain_repeats, batched_meshes=args.batched_meshes,
                                        batched_workers=args.batched_workers))
    reducer = leg(args.no_reducer, "reducer", lambda: reducer_leg(3))
    executor = leg(args.no_executor, "executor",
                   lambda: executor_leg(args.executor_budget, 0 if args.no_cpu else args.executor_cpu_budget))

    line = {
        "metric": METRIC,
        "value": round(1000.0 / ms_per_step, 3),
        "unit": "audits/s (64 views x 1024^2 samples, reference + candidate rasters, EDT Hausdorff, changed area)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded fBm heightfield and its quadric-collapsed copy)",
        "config": {"workload": "raster audit alone: 512x512 fBm heightfield (262,144 V / 522,242 F) vs "
                               f"25%-collapsed copy ({cand.triangles} F), 64 ortho views at 1024x1024 samples "
                               "(S=120, ss=8), coverage profile, limit 1e3",
                   "views": views, "samples_per_view": samples, "reference_triangles": ref.triangles,
                   "candidate_triangles": cand.triangles},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(),
                     "algorithmic_bytes_per_launch": per_launch_bytes,
                     "kernel_ms_per_launch": {k: round(v, 4) for k, v in kernels.items()},
                     "raster_pass_ms": round(raster_ms / launches, 4),
                     "face_split_ms": round(split_ms / band_launches, 4),
                     "raster_chunks_per_view": audit_chunks,
                     "fp64_vector": {"kernel": "raster_band_kernel", "ops_per_launch": fp64_ops,
                                     "achieved": round(fp64_tflops, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                     "frac": round(fp64_tflops / FP64_PEAK_TFLOPS, 4)}},
        "audit_roofline": {"algorithmic_bytes": b["audit"],
                           "achieved_gbs": round(b["audit"] / (ms_per_step * 1e-3) / 1e9, 2),
                           "frac": round(b["audit"] / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)},
        "measurement": m.as_tuple() if m else None,
        "cpu_baseline": cpu,
        "parity_on_cpu_sample": parity,
        "chain": chain,
        "reducer": reducer,
        "patches": patches,
        "executor": executor,
        "appearance": appearance,
        "neural": neural,
        "training": training,
        "scorer": scorer,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
