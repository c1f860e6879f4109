"""Loader for the fixtures written by tests/golden/make_golden.py."""
from __future__ import annotations

import hashlib
import json
from functools import lru_cache
from pathlib import Path

import numpy as np

from blitzremesher_amd import _lib
from blitzremesher_amd.audit import EvalSettings, Measurement, Mesh

HERE = Path(__file__).resolve().parent / "golden"


@lru_cache(maxsize=1)
def audit_golden() -> dict:
    return json.loads((HERE / "audit_golden.json").read_text())


@lru_cache(maxsize=1)
def golden_meshes() -> dict[str, Mesh]:
    data = np.load(HERE / "meshes.npz")
    names = sorted({k.split("/")[0] for k in data.files})
    out = {}
    for n in names:
        opt = {k: data[f"{n}/{k}"] if f"{n}/{k}" in data.files else None
               for k in ("double_sided", "materials", "normals", "colors")}
        out[n] = Mesh(data[f"{n}/positions"], data[f"{n}/indices"], **opt)
    return out


def camera_of(d: dict) -> _lib.Camera:
    c = _lib.Camera()
    for i in range(3):
        c.right[i], c.up[i], c.forward[i] = d["right"][i], d["up"][i], d["forward"][i]
    c.distance, c.focal, c.scale, c.perspective = d["distance"], d["focal"], d["scale"], d["perspective"]
    return c


def camera_fields(c) -> dict:
    return {"right": list(c.right), "up": list(c.up), "forward": list(c.forward), "distance": c.distance,
            "focal": c.focal, "scale": c.scale, "perspective": int(c.perspective)}


def settings_of(d: dict) -> EvalSettings:
    return EvalSettings(**d)


def measurement_of(d: dict) -> Measurement:
    return Measurement(**d)


def digest(words) -> str:
    return hashlib.sha256(np.ascontiguousarray(words, np.uint64).tobytes()).hexdigest()


def pixel_digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def unpack_mask(bits: list[int], size: int) -> np.ndarray:
    return np.unpackbits(np.array(bits, np.uint8))[: size * size].reshape(size, size).astype(bool)


def nonfinite_meshes():
    """Meshes whose position streams hold NaN or infinite coordinates (the
    reference's evaluate does not reject them), with the bounds of their
    finite part: name -> (mesh, center, radius). A NaN first corner makes the
    bounding box non-finite (the face counts as out of frame); a NaN second
    or third corner leaves a finite box whose edge tests are NaN (every
    sample of the box passes); an infinite corner is out of frame."""
    import numpy as np
    from blitzremesher_amd import Mesh, meshes
    g = meshes.grid(6)
    out = {}
    for name, vertex, value in (("nan_first", None, np.nan), ("nan_inner", 20, np.nan), ("inf", 9, np.inf),
                                ("neg_inf", 30, -np.inf)):
        pos = g.positions.copy()
        idx = g.indices.copy()
        if vertex is None:
            vertex = int(idx[len(idx) // 2, 0])  # the first corner of a middle face
        pos[vertex, 0] = value
        finite = np.isfinite(pos).all(1)
        lo, hi = pos[finite].min(0), pos[finite].max(0)
        center = ((lo + hi) * .5).astype(np.float32)
        radius = float(np.linalg.norm(hi - lo) * .5)
        out[name] = (Mesh(pos, idx), center, radius)
    return out
