"""Benchmark of the HIP coverage audit on the raster-audit workload.

Workload (one step): a full coverage audit of a seeded 512 x 512 fBm
heightfield (262,144 V / 522,242 F) against its 25 %-collapsed copy (quadric
collapses with reused vertices, 391,681 F) from 64 orthographic Fibonacci
views at 1024 x 1024 samples (screen 120 px, supersample 8), limit 1e3 so
every view is measured. Meshes are resident in HBM before timing starts.

Prints one JSON line with the throughput, the dominant kernel's roofline
(HIP-event kernel times from the library's own stream, algorithmic bytes as
documented in DESIGN.md) and a CPU baseline timed on a bounded sample of the
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
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "audited collapses/s & ms per LOD chain at ≤1px error; raster audit % HBM roof"


def workload():
    from blitzremesher_amd import EvalSettings, meshes
    from blitzremesher_amd.audit import bounds, reduce_reuse
    ref = meshes.heightfield(512, 0.1, 2)
    cand = reduce_reuse(ref, int(round(ref.triangles * 0.75)))
    center, radius = bounds(ref)
    # The heightfield is an open surface seen from all around the sphere, so
    # both sides are drawn (force_two_sided); with one-sided faces the views
    # from below are empty or near-empty and the audit stops there.
    settings = EvalSettings(profile=0, orthographic=64, perspective=0, rotation_seed=0xB1172026, screen_size=120,
                            supersample=8, max_supersample=8, limit=1e3, max_changed_area=1.0,
                            force_two_sided=True)
    return ref, cand, center, radius, settings


def algorithmic_bytes(ref, cand, views: int, samples: int, distance_views: int) -> dict:
    """Bytes an ideal audit must move ('Roofline accounting'):
    geometry read once (12 B per vertex and per face of each mesh), one bit
    per sample written for each of the two masks per view, and for each view
    needing the distance transform two directions of 16-bit row distances
    written and read (8 B per sample)."""
    geometry = 12 * (ref.positions.shape[0] + ref.triangles) + 12 * (cand.positions.shape[0] + cand.triangles)
    masks = views * 2 * samples // 8
    return {"geometry": geometry, "masks": masks, "raster": geometry + masks,
            "distance": distance_views * 8 * samples, "audit": geometry + 2 * masks + distance_views * 8 * samples}


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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-sample-views", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} (WORLD_SIZE {world}) is not supported: every box here "
              f"has one GPU and the audit path has no multi-GPU leg", file=sys.stderr)
        return 2

    from blitzremesher_amd import Audit, _lib
    ref, cand, center, radius, settings = workload()
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
        raster_ms = compare_ms = distance_ms = 0.0
        launches = 0
        distance_views = 0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            m = audit.evaluate(dref, dcand, center, radius, settings)
            t = audit.last_timing()
            raster_ms += t["raster_ms"]
            compare_ms += t["compare_ms"]
            distance_ms += t["distance_ms"]
            launches += t["passes"]
            distance_views = t["distance_views"]
        sync()
        elapsed = time.perf_counter() - t0
        ms_per_step = 1000.0 * elapsed / args.steps
        samples = (int(-(-(settings.screen_size + 8) // 1)) * settings.supersample) ** 2
        views = settings.orthographic + settings.perspective
        b = algorithmic_bytes(ref, cand, views, samples, views)
        kernels = {"raster": raster_ms / launches, "compare": compare_ms / launches,
                   "distance": distance_ms / launches}
        dominant = max(kernels, key=kernels.get)
        per_launch_bytes = {"raster": b["raster"], "compare": 2 * b["masks"], "distance": b["distance"]}[dominant]
        achieved = per_launch_bytes / (kernels[dominant] * 1e-3) / 1e9
        gpu_sample = None
        cpu = None
        parity = None
        if not args.no_cpu:
            cpu, sample, cpu_m = cpu_baseline(ref, cand, center, radius, settings, args.cpu_sample_views)
            gpu_sample = audit.evaluate(dref, dcand, center, radius, sample)
            parity = gpu_sample.as_tuple() == cpu_m.as_tuple()

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
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "algorithmic_bytes_per_launch": per_launch_bytes,
                     "kernel_ms_per_launch": {k: round(v, 4) for k, v in kernels.items()}},
        "audit_roofline": {"algorithmic_bytes": b["audit"],
                           "achieved_gbs": round(b["audit"] / (ms_per_step * 1e-3) / 1e9, 2),
                           "frac": round(b["audit"] / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)},
        "measurement": m.as_tuple() if m else None,
        "cpu_baseline": cpu,
        "parity_on_cpu_sample": parity,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
