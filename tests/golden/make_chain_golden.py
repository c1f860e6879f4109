"""Writes tests/golden/chain_golden.json and chain_meshes.npz from the
reference's own chain search (detail::generate_with_hooks, src/chain.cpp),
built into oracle/_ref/libblitz_ref.so by oracle/Makefile, driven by the CPU
hooks of oracle/ref_shim.cpp (ref_generate_hooks). Run in the build container:

    make -f oracle/Makefile ref && python tests/golden/make_chain_golden.py

Fixtures are data: the meshes, and per case the per-level triangle counts,
index-buffer digests, schedule, both audit records of every level, the
generation counters and the finalists' triangle counts.
"""
from __future__ import annotations

import hashlib
import json
import sys
from concurrent.futures import ProcessPoolExecutor
from dataclasses import asdict
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE.parent))

from chain_cases import chain_cases  # noqa: E402
from oracle_lib import ref_generate_hooks  # noqa: E402


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.uint32).tobytes()).hexdigest()


def run(name: str) -> dict:
    mesh, settings, options = chain_cases()[name]
    r = ref_generate_hooks(mesh, (settings, options))
    return {
        "case": name,
        "triangles": [int(l.shape[0]) for l in r["lods"]],
        "indices_sha256": [sha(l) for l in r["lods"]],
        "schedule": r["schedule"],
        "source": [list(m.as_tuple()) for m in r["source"]],
        "adjacent": [list(m.as_tuple()) for m in r["adjacent"]],
        "counters": {k: r[k] for k in ("status", "candidate_evaluations", "audit_evaluations", "rejected_gates",
                                       "legal_collapses", "rejected_collapses", "action_trials", "action_ranked",
                                       "decoded", "finalists", "fallback_levels")},
        "finalist_triangles": r["finalist_triangles"],
    }


def main() -> None:
    cases = chain_cases()
    arrays = {}
    for name, (mesh, _, _) in cases.items():
        arrays[f"{name}/positions"] = mesh.positions
        arrays[f"{name}/indices"] = mesh.indices
    with ProcessPoolExecutor(len(cases)) as pool:
        results = list(pool.map(run, list(cases)))
    np.savez_compressed(HERE / "chain_meshes.npz", **arrays)
    out = {"cases": results}
    (HERE / "chain_golden.json").write_text(json.dumps(out, indent=1) + "\n")
    for r in results:
        print(r["case"], r["triangles"], r["counters"]["legal_collapses"])


if __name__ == "__main__":
    main()
