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
