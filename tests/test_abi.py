"""CPU checks of the C ABI: the library loads, exports every declared symbol,
and its host-side pieces (bounds, cameras, descriptor validation, defaults)
agree with the oracle. No kernel is launched."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from blitzremesher_amd import _lib, audit, meshes
from oracle_lib import Oracle

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols() -> set[str]:
    names = set()
    for header in (ROOT / "include").glob("*.h"):
        text = header.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(blitz_hip_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.PROTOTYPES) >= {n for n in names if n.startswith("blitz_hip_") and n in _lib.PROTOTYPES}


def test_api_version_and_defaults():
    lib = _lib.load()
    assert lib.blitz_hip_api_version() == 1
    s = _lib.EvalSettingsC()
    assert lib.blitz_hip_eval_settings_init(C.byref(s), C.sizeof(s)) == 0
    # include/blitz/evaluate.hpp:24-33 defaults.
    assert (s.profile, s.supersample, s.max_supersample) == (1, 8, 32)
    assert (s.views_orthographic, s.views_perspective, s.views_rotation_seed) == (642, 64, 0xB1172026)
    assert (s.screen_size, s.limit, s.max_changed_area) == (512.0, 2.0, 1.0)
    assert lib.blitz_hip_eval_settings_init(C.byref(s), C.sizeof(s) - 1) == 1


@pytest.mark.parametrize("kind", ["oracle", "ref"])
def test_bounds_and_cameras_match(kind):
    try:
        o = Oracle(kind)
    except RuntimeError:
        pytest.skip(f"{kind} library not built")
    for mesh in (meshes.grid(8), meshes.icosphere(2, 0.02, 1), meshes.small_torus()):
        c, r = audit.bounds(mesh)
        oc, orad = o.bounds(mesh)
        assert np.array_equal(c, oc) and r == orad
        for screen, ortho, persp, seed in ((17.25, 8, 4, 17), (256, 16, 0, 0xB1172026), (4000, 3, 5, 1)):
            got = audit.cameras(c, r, screen, ortho, persp, seed)
            want = o.cameras(c, r, screen, ortho, persp, seed)
            for a, b in zip(got, want):
                for f, _ in a._fields_:
                    x, y = getattr(a, f), getattr(b, f)
                    assert (list(x) if hasattr(x, "__len__") else x) == (list(y) if hasattr(y, "__len__") else y)


def test_descriptor_validation_is_loud():
    lib = _lib.load()
    m = meshes.grid(2)
    d = m.descriptor()
    b = _lib.Bounds()
    err = _lib.error_buffer()
    assert lib.blitz_hip_bounds_of(C.byref(d), C.byref(b), err, len(err)) == 0
    d.abi_version = 4
    assert lib.blitz_hip_bounds_of(C.byref(d), C.byref(b), err, len(err)) == 1
    assert b"ABI" in err.value
    bad = audit.Mesh(m.positions, np.array([[0, 1, 99]], np.uint32))
    d = bad.descriptor()
    assert lib.blitz_hip_bounds_of(C.byref(d), C.byref(b), err, len(err)) == 1
    assert b"index outside" in err.value


def test_create_without_device_is_unavailable():
    if _lib.load().blitz_hip_available(0):
        pytest.skip("a device is present")
    with pytest.raises(_lib.BlitzError) as e:
        audit.Audit(0)
    assert e.value.status == 6
