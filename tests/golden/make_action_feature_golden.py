"""Writes action_features.npz: the reference build's graph channels
(neural::graph, src/neural/graph.cpp:35-187) and action feature rows
(ActionState::actions, src/neural/action.cpp:371-450) for the meshes of
tests/feature_cases.py, with the meshes themselves. Run from the repository
root after `make -C oracle ref`."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
from feature_cases import CONDITIONS, cases  # noqa: E402
from oracle_lib import Actions  # noqa: E402


def main() -> None:
    out = {"conditions": CONDITIONS}
    for name, (mesh, uv, tangents) in cases().items():
        ref = Actions("ref", mesh, uv, tangents)
        out[f"{name}/positions"] = mesh.positions
        out[f"{name}/indices"] = mesh.indices
        for key, value in (("materials", mesh.materials), ("normals", mesh.normals), ("colors", mesh.colors),
                           ("uv", uv), ("tangents", tangents)):
            if value is not None:
                out[f"{name}/{key}"] = value
        out[f"{name}/legal"] = ref.legal
        out[f"{name}/graph"] = ref.ref_graph_features()
        out[f"{name}/rows"] = ref.ref_features(CONDITIONS)
        print(name, mesh.positions.shape[0], "vertices", len(ref.legal), "legal actions")
        ref.close()
    np.savez_compressed(HERE / "action_features.npz", **out)


if __name__ == "__main__":
    main()
