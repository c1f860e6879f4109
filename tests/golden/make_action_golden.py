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

from action_meshes import action_cases  # noqa: E402
from oracle_lib import Actions  # noqa: E402

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
    np.savez_compressed(HERE / "action_meshes.npz", **arrays)
    (HERE / "action_golden.json").write_text(json.dumps(
        {"source": "reference ActionState (src/neural/action.cpp) built by oracle/Makefile", "cases": cases},
        indent=1))
    print(f"wrote {len(cases)} action cases")


if __name__ == "__main__":
    main()
