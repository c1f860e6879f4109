"""Writes the golden fixtures of tests/golden/ from a build of the reference's
own CPU sources (oracle/_ref/libblitz_ref.so, built by oracle/Makefile from the
reference checkout). Run in the build container only:

    make -f oracle/Makefile ref && python tests/golden/make_golden.py

Fixtures are data: input meshes, settings and the reference's outputs
(cameras, mask digests, distances, Measurements, reduced index buffers).
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

from blitzremesher_amd import meshes  # noqa: E402
from blitzremesher_amd.audit import EvalSettings, Mesh  # noqa: E402
from oracle_lib import Oracle, mask_to_words  # noqa: E402


def digest(words: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(words, np.uint64).tobytes()).hexdigest()


def pixel_digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def camera_dict(c) -> dict:
    return {"right": list(c.right), "up": list(c.up), "forward": list(c.forward), "distance": c.distance,
            "focal": c.focal, "scale": c.scale, "perspective": int(c.perspective)}


def fixture_meshes(ref: Oracle) -> dict[str, Mesh]:
    out: dict[str, Mesh] = {}
    g4 = meshes.grid(4)
    out["grid4"] = g4
    idx = np.concatenate([g4.indices, np.zeros((1, 3), np.uint32)])
    pos = g4.positions.copy()
    pos[:, 2] += np.float32(1e-5)
    out["grid4_one_sided_degenerate"] = Mesh(pos, idx, double_sided=np.array([0], np.uint8))
    pos2 = pos.copy()
    pos2[:, 0] += np.float32(100)
    out["grid4_offscreen"] = Mesh(pos2, idx, double_sided=np.array([0], np.uint8))
    out["grid8"] = meshes.grid(8)
    out["grid8_moved"] = Mesh(out["grid8"].positions + np.array([2, 0, 0], np.float32), out["grid8"].indices,
                              double_sided=out["grid8"].double_sided)
    tiles = meshes.four_tiles()
    out["tiles"] = tiles
    out["tiles_minus_first"] = Mesh(tiles.positions, tiles.indices[2:], double_sided=tiles.double_sided)
    out["tiles_minus_last"] = Mesh(tiles.positions, tiles.indices[:6], double_sided=tiles.double_sided)
    out["tiles_empty"] = Mesh(tiles.positions, np.zeros((0, 3), np.uint32), double_sided=tiles.double_sided)
    out["torus_small"] = meshes.small_torus()
    ico = meshes.icosphere(3, 0.02, 1)
    out["ico3"] = ico
    out["ico3_jitter"] = meshes.jitter(ico, 0.004, 5, 0.3)
    out["ico3_qem_900"] = ref.reduce_reuse(ico, 900)
    out["ico3_qem_300"] = ref.reduce_reuse(ico, 300)
    t = meshes.torus(1.0, 0.3, 60, 24, 0.01, 3)
    out["torus60"] = t
    out["torus60_qem_1500"] = ref.reduce_reuse(t, 1500)
    g24 = meshes.grid(24)
    out["grid24"] = g24
    out["grid24_cut"] = Mesh(g24.positions, g24.indices[: len(g24.indices) * 2 // 3], double_sided=g24.double_sided)
    hf = meshes.heightfield(48, 0.1, 2)
    out["heightfield48"] = hf
    out["heightfield48_qem"] = ref.reduce_reuse(hf, int(hf.triangles * 0.75))
    # Appearance streams (blitz::MeshView normals/colours, per-face materials).
    out["ico2_attr"] = meshes.with_attributes(meshes.icosphere(2, 0.02, 1), 7)
    out["ico2_attr_jitter"] = meshes.with_attributes(meshes.jitter(meshes.icosphere(2, 0.02, 1), 0.01, 3, 0.4), 7)
    hfa = meshes.with_attributes(meshes.heightfield(20, 0.1, 2), 9, colors=True)
    out["hf20_attr"] = hfa
    out["hf20_attr_materials"] = Mesh(hfa.positions, hfa.indices, (np.arange(hfa.triangles) % 3).astype(np.uint16),
                                      np.array([0, 1, 0], np.uint8), hfa.normals, hfa.colors)
    out["hf20_attr_relabelled"] = Mesh(hfa.positions, hfa.indices, (np.arange(hfa.triangles) % 3).astype(np.uint16),
                                       np.zeros(3, np.uint8), hfa.normals, hfa.colors)
    out["ico2_plain"] = meshes.icosphere(2, 0.02, 1)
    return out


def settings_dict(s: EvalSettings) -> dict:
    return {k: getattr(s, k) for k in ("profile", "supersample", "max_supersample", "screen_size", "limit",
                                       "max_changed_area", "orthographic", "perspective", "rotation_seed",
                                       "force_two_sided", "conservative_screen", "normal_weight", "color_weight",
                                       "material_weight")}


def main() -> None:
    ref = Oracle("ref")
    ms = fixture_meshes(ref)
    arrays = {}
    for name, m in ms.items():
        arrays[f"{name}/positions"] = m.positions
        arrays[f"{name}/indices"] = m.indices
        for field in ("double_sided", "materials", "normals", "colors"):
            if getattr(m, field) is not None:
                arrays[f"{name}/{field}"] = getattr(m, field)
    np.savez_compressed(HERE / "meshes.npz", **arrays)

    # Cameras (src/evaluate.cpp:108-124).
    cams = []
    for name, screen, ortho, persp, seed in [("grid4", 17.25, 8, 4, 17), ("grid4", 17.25, 8, 4, 9381),
                                             ("ico3", 256, 16, 0, 0xB1172026), ("torus60", 120, 64, 0, 0xB1172026),
                                             ("grid8", 4000, 3, 5, 1), ("tiles", 64, 4, 0, 460796)]:
        c, r = ref.bounds(ms[name])
        cams.append({"mesh": name, "center": [float(x) for x in c], "radius": r, "screen": screen, "ortho": ortho,
                     "persp": persp, "seed": seed,
                     "cameras": [camera_dict(k) for k in ref.cameras(c, r, screen, ortho, persp, seed)]})

    # Coverage masks (src/evaluate.cpp:125-189), tests/precision.cpp shapes.
    rasters = []

    def add_raster(name, mesh_name, center, radius, cam, screen, ss, two):
        r = ref.raster(ms[mesh_name], center, radius, cam, screen, ss, two)
        entry = {"case": name, "mesh": mesh_name, "center": [float(x) for x in center], "radius": radius,
                 "camera": camera_dict(cam), "screen": screen, "ss": ss, "two_sided": two}
        if r is None:
            entry["limited"] = True
        else:
            words, size, clipped = r
            entry.update({"size": size, "clipped": clipped, "sha256": digest(words),
                          "popcount": int(sum(bin(int(w)).count("1") for w in words))})
        rasters.append(entry)

    c, r = ref.bounds(ms["grid4"])
    for seed in (17, 9381):
        for ss in (1, 2, 4, 8):
            for cam in ref.cameras(c, r, 17.25, 8, 4, seed):
                for two in (False, True):
                    add_raster("grid4_views", "grid4", c, r, cam, 17.25, ss, two)
    persp = ref.cameras(c, r, 17.25, 1, 1, 99)[-1]
    add_raster("cull_degenerate", "grid4_one_sided_degenerate", c, r, persp, 17.25, 2, False)
    add_raster("offscreen", "grid4_offscreen", c, r, persp, 17.25, 2, True)
    persp.distance = -100.0
    add_raster("behind_camera", "grid4_offscreen", c, r, persp, 17.25, 2, True)
    for mesh_name in ("torus60", "heightfield48"):
        c2, r2 = ref.bounds(ms[mesh_name])
        for screen in (23, 24, 25, 57, 58, 59, 121):
            for cam in ref.cameras(c2, r2, screen, 2, 1, 0xB1172026):
                add_raster("tile_straddle", mesh_name, c2, r2, cam, screen, 2, False)

    # Coverage distance (src/evaluate.cpp:196-216) on seeded random masks.
    rng = np.random.default_rng(871)
    distances = []
    for trial in range(40):
        size = int(rng.integers(1, 40))
        a = rng.random((size, size)) < 0.2
        b = rng.random((size, size)) < 1 / 7
        for ss in (1, 2, 8):
            distances.append({"size": size, "ss": ss, "a": np.packbits(a.reshape(-1)).tolist(),
                              "b": np.packbits(b.reshape(-1)).tolist(),
                              "value": ref.coverage_distance(mask_to_words(a), mask_to_words(b), size, ss)})

    # Full audits (src/evaluate.cpp:312-352).
    evals = []

    def add_eval(name, ref_name, cand_name, s: EvalSettings, bounds_of=None):
        c, r = ref.bounds(ms[bounds_of or ref_name])
        m = ref.evaluate(ms[ref_name], ms[cand_name], c, r, s)
        evals.append({"case": name, "reference": ref_name, "candidate": cand_name, "center": [float(x) for x in c],
                      "radius": r, "settings": settings_dict(s),
                      "measurement": {k: (bool(v) if isinstance(v, bool) else v) for k, v in m.__dict__.items()}})

    base = dict(profile=0, screen_size=40, supersample=2, max_supersample=8, orthographic=6, perspective=2,
                rotation_seed=0xB1172026)
    for limit in (0.25, 1.0, 3.0, 1000.0):
        add_eval("ico_jitter", "ico3", "ico3_jitter", EvalSettings(**base, limit=limit))
        add_eval("ico_qem900", "ico3", "ico3_qem_900", EvalSettings(**base, limit=limit))
        add_eval("ico_qem300", "ico3", "ico3_qem_300", EvalSettings(**base, limit=limit))
        add_eval("torus_qem", "torus60", "torus60_qem_1500", EvalSettings(**base, limit=limit))
        add_eval("heightfield_qem", "heightfield48", "heightfield48_qem", EvalSettings(**base, limit=limit))
    for ss, mx, limit in [(1, 8, 0.6), (2, 16, 0.8), (3, 12, 1.0), (4, 4, 1.0), (1, 32, 0.5)]:
        b2 = dict(base, supersample=ss, max_supersample=mx, screen_size=33)
        add_eval("ico_refine", "ico3", "ico3_qem_900", EvalSettings(**b2, limit=limit))
    contract = dict(profile=0, screen_size=24, orthographic=8, perspective=2, rotation_seed=1, supersample=4,
                    max_supersample=8)
    add_eval("identical", "grid8", "grid8", EvalSettings(**contract))
    add_eval("moved", "grid8", "grid8_moved", EvalSettings(**contract))
    add_eval("limited", "grid8", "grid8_moved", EvalSettings(**dict(contract, screen_size=4096, supersample=8,
                                                                      max_supersample=8)))
    for views in (1, 4):
        for mca in (1.0, 0.35, 0.1, 0.0):
            s = EvalSettings(profile=0, orthographic=views, perspective=0, rotation_seed=460796, screen_size=64,
                             supersample=4, max_supersample=4, force_two_sided=True, limit=1000,
                             max_changed_area=mca, normal_weight=0, color_weight=0, material_weight=0)
            for a, b in [("tiles", "tiles_minus_first"), ("tiles", "tiles_minus_last"),
                         ("tiles_minus_first", "tiles_minus_last"), ("tiles", "tiles_empty"),
                         ("tiles_empty", "tiles")]:
                add_eval(f"tiles_v{views}_a{mca}", a, b, s, bounds_of="tiles")
    for screen in (23, 24, 25, 57, 58, 59, 121):
        add_eval("cut_pattern", "grid24", "grid24_cut",
                 EvalSettings(**dict(base, screen_size=screen, orthographic=3, perspective=1, max_supersample=4),
                              limit=2.0))
    for conservative in (False, True):
        add_eval("torus_conservative", "torus60", "torus60_qem_1500",
                 EvalSettings(**dict(base, orthographic=2, perspective=6), conservative_screen=conservative,
                              limit=1.5))

    # Attribute rasters (src/evaluate.cpp:218-280): per-stream digests.
    pixels = []
    for mesh_name in ("ico2_attr", "hf20_attr_materials"):
        c, r = ref.bounds(ms[mesh_name])
        for cam in ref.cameras(c, r, 24, 4, 3, 77):
            for ss in (1, 3):
                for two in (False, True):
                    got = ref.raster_pixels(ms[mesh_name], c, r, cam, 24, ss, two)
                    entry = {"mesh": mesh_name, "center": [float(x) for x in c], "radius": r,
                             "camera": camera_dict(cam), "screen": 24, "ss": ss, "two_sided": two}
                    if got is None:
                        entry["limited"] = True
                    else:
                        streams, size, clipped = got
                        entry.update({"size": size, "clipped": clipped,
                                      "sha256": {k: pixel_digest(v) for k, v in streams.items()},
                                      "visible": int(streams["visible"].sum())})
                    pixels.append(entry)

    # Normals and attributes profiles (src/evaluate.cpp:282-352).
    attributed = []
    for profile in (1, 2):
        for ref_name, cand_name in [("ico2_attr", "ico2_attr_jitter"), ("hf20_attr", "hf20_attr_materials"),
                                    ("hf20_attr", "hf20_attr_relabelled"), ("ico2_plain", "ico2_attr"), ("ico2_attr", "ico2_attr")]:
            for limit in (0.5, 2.0, 40.0):
                for ss, mx in ((1, 4), (2, 2)):
                    s = EvalSettings(profile=profile, screen_size=20, supersample=ss, max_supersample=mx,
                                     orthographic=6, perspective=3, limit=limit, max_changed_area=1.0)
                    add_eval(f"profile{profile}", ref_name, cand_name, s)
                    attributed.append(evals.pop())

    (HERE / "audit_golden.json").write_text(json.dumps(
        {"source": "reference CPU sources (src/evaluate.cpp, src/mesh.cpp, src/simplify.cpp) built by oracle/Makefile",
         "cameras": cams, "rasters": rasters, "distances": distances, "evaluations": evals, "pixels": pixels,
         "attributed": attributed}, indent=1))
    print(f"wrote {len(cams)} camera sets, {len(rasters)} rasters, {len(distances)} distances, "
          f"{len(evals)} evaluations, {len(pixels)} attribute rasters, {len(attributed)} attributed evaluations")


if __name__ == "__main__":
    main()
