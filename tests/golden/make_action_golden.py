"""Writes tests/golden/action_golden.json and action_meshes.npz from a build of
the reference's own ActionState (src/neural/action.cpp, compiled into
oracle/_ref/libblitz_ref.so by oracle/Makefile). Run in the build container:

    make -f oracle/Makefile ref && python tests/golden/make_action_golden.py

Fixtures are data: the meshes (and UVs), and per mesh the legal directed
collapses in enumeration order, independent batches for seeded rank orders
and the trial index buffers of those batches.
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE.parent))

from action_meshes import action_cases, executor_cases  # noqa: E402
from oracle_lib import Actions, Oracle, ref_execute_actions  # noqa: E402

BATCHES = ((64, 10 ** 9, 11), (5, 10 ** 9, 12), (64, 7, 13), (1, 10 ** 9, 14))


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.uint32).tobytes()).hexdigest()


def main() -> None:
    arrays, cases = {}, []
    for name, (mesh, uv) in action_cases().items():
        arrays[f"{name}/positions"] = mesh.positions
        arrays[f"{name}/indices"] = mesh.indices
        if mesh.materials is not None:
            arrays[f"{name}/materials"] = mesh.materials
        if uv is not None:
            arrays[f"{name}/uv"] = uv
        ref = Actions("ref", mesh, uv)
        entry = {"mesh": name, "legal": int(len(ref.legal)), "legal_sha256": sha(ref.legal), "batches": []}
        for maximum, budget, seed in BATCHES:
            order = np.random.default_rng(seed).permutation(len(ref.legal)).astype(np.uint32)
            picked = ref.independent(order, maximum, budget)
            b = {"seed": seed, "maximum": maximum, "maximum_removed": budget, "selected": picked.tolist()}
            if len(picked):
                t = ref.trial(picked)
                b.update({"trial_faces": int(len(t)), "trial_sha256": sha(t)})
            entry["batches"].append(b)
        cases.append(entry)
        ref.close()
    # neural::execute_actions with the reference CPU audit as the gate.
    executions = []
    ref = Oracle("ref")
    for name, mesh, gate, target, budget in executor_cases():
        arrays[f"exec_{name}/positions"] = mesh.positions
        arrays[f"exec_{name}/indices"] = mesh.indices
        if mesh.double_sided is not None:
            arrays[f"exec_{name}/double_sided"] = mesh.double_sided
        center, radius = ref.bounds(mesh)
        for ranking in (1, 3, 4):
            for batch in (1, 4, 16):
                out, stats = ref_execute_actions(mesh, center, radius, gate, target, budget, ranking, batch)
                executions.append({"mesh": f"exec_{name}", "center": [float(x) for x in center], "radius": radius,
                                   "gate": {k: v for k, v in gate.__dict__.items() if k != "cancelled"},
                                   "target": target, "budget": budget, "ranking": ranking, "batch": batch,
                                   "faces": int(len(out)), "indices_sha256": sha(out), "stats": stats})
    np.savez_compressed(HERE / "action_meshes.npz", **arrays)
    (HERE / "action_golden.json").write_text(json.dumps(
        {"source": "reference ActionState and execute_actions (src/neural/action.cpp) built by oracle/Makefile",
         "cases": cases, "executions": executions}, indent=1))
    print(f"wrote {len(cases)} action cases, {len(executions)} executions")


if __name__ == "__main__":
    main()
