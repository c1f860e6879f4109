"""Python view of the HIP coverage audit (include/blitz_hip.h).

Mirrors the reference's C++ audit interface for tests and benchmarks:
``EvalSettings`` and ``Measurement`` follow include/blitz/evaluate.hpp:24-39,
``Audit.evaluate`` is ``evaluate_gpu`` / ``AuditCuda::evaluate``
(include/blitz/neural.hpp:148-149), ``Audit.raster_coverage`` is
``detail::rasterize_coverage`` (src/coverage.hpp:12) and
``Audit.coverage_distance`` is ``detail::coverage_distance`` (src/coverage.hpp:13).
All work runs in libblitz_hip.so on a gfx950 device.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field, fields

import numpy as np

from . import _lib
from ._lib import BlitzError

PROFILE_COVERAGE, PROFILE_NORMALS, PROFILE_ATTRIBUTES = 0, 1, 2


@dataclass
class Mesh:
    """Positions (n, 3) float32 and triangle indices (m, 3) uint32, with the
    optional per-face materials, per-material two-sided flags, per-vertex
    normals (n, 3) float32 and RGBA8 colours (n, 4) uint8 of blitz::MeshView
    (include/blitz/mesh.hpp:91-105)."""
    positions: np.ndarray
    indices: np.ndarray
    materials: np.ndarray | None = None
    double_sided: np.ndarray | None = None
    normals: np.ndarray | None = None
    colors: np.ndarray | None = None

    def __post_init__(self) -> None:
        self.positions = np.ascontiguousarray(self.positions, dtype=np.float32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(self.indices, dtype=np.uint32).reshape(-1, 3)
        if self.materials is not None:
            self.materials = np.ascontiguousarray(self.materials, dtype=np.uint16).reshape(-1)
        if self.double_sided is not None:
            self.double_sided = np.ascontiguousarray(self.double_sided, dtype=np.uint8).reshape(-1)
        if self.normals is not None:
            self.normals = np.ascontiguousarray(self.normals, dtype=np.float32).reshape(-1, 3)
        if self.colors is not None:
            self.colors = np.ascontiguousarray(self.colors, dtype=np.uint8).reshape(-1, 4)

    @property
    def triangles(self) -> int:
        return int(self.indices.shape[0])

    def descriptor(self) -> _lib.BlitzMesh:
        """ABI-5 blitz_mesh borrowing this mesh's arrays."""
        d = _lib.BlitzMesh()
        d.struct_size = C.sizeof(_lib.BlitzMesh)
        d.abi_version = _lib.BLITZ_ABI_VERSION
        n = self.positions.shape[0]
        d.positions = _lib.BlitzStream(self.positions.ctypes.data if n else None, n, 12)
        d.indices = self.indices.ctypes.data_as(C.POINTER(C.c_uint32)) if self.indices.size else None
        d.index_count = int(self.indices.size)
        if self.materials is not None and self.materials.size:
            d.materials = self.materials.ctypes.data_as(C.POINTER(C.c_uint16))
            d.material_count = int(self.materials.size)
        if self.double_sided is not None and self.double_sided.size:
            d.double_sided = self.double_sided.ctypes.data_as(C.POINTER(C.c_uint8))
            d.double_sided_count = int(self.double_sided.size)
        if self.normals is not None and n:
            d.normals = _lib.BlitzStream(self.normals.ctypes.data, n, 12)
        if self.colors is not None and n:
            d.colors = _lib.BlitzStream(self.colors.ctypes.data, n, 4)
        return d


@dataclass
class EvalSettings:
    """blitz::EvalSettings defaults (include/blitz/evaluate.hpp:24-33)."""
    profile: int = PROFILE_NORMALS
    supersample: int = 8
    max_supersample: int = 32
    screen_size: float = 512.0
    limit: float = 2.0
    max_changed_area: float = 1.0
    orthographic: int = 642
    perspective: int = 64
    rotation_seed: int = 0xB1172026
    force_two_sided: bool = False
    conservative_screen: bool = False
    normal_weight: float = 180.0 / (10 * math.pi)
    color_weight: float = 4.0
    material_weight: float = 4.0
    cancelled: object | None = None  # callable() -> bool, polled once per view

    def to_c(self) -> tuple[_lib.EvalSettingsC, object]:
        s = _lib.EvalSettingsC()
        s.struct_size = C.sizeof(_lib.EvalSettingsC)
        s.version = _lib.BLITZ_HIP_API_VERSION
        s.profile = self.profile
        s.supersample = self.supersample
        s.max_supersample = self.max_supersample
        s.force_two_sided = int(self.force_two_sided)
        s.conservative_screen = int(self.conservative_screen)
        s.views_orthographic = self.orthographic
        s.views_perspective = self.perspective
        s.views_rotation_seed = self.rotation_seed & 0xFFFFFFFF
        s.screen_size = self.screen_size
        s.limit = self.limit
        s.max_changed_area = self.max_changed_area
        s.normal_weight = self.normal_weight
        s.color_weight = self.color_weight
        s.material_weight = self.material_weight
        keep = None
        if self.cancelled is not None:
            fn = self.cancelled
            keep = _lib.CANCEL_FN(lambda _user: 1 if fn() else 0)
            s.cancelled = keep
        return s, keep


@dataclass
class Measurement:
    """blitz::Measurement (include/blitz/evaluate.hpp:34-39)."""
    error: float = 0.0
    coverage: float = 0.0
    coverage_upper: float = 0.0
    changed_area: float = 0.0
    normal_degrees: float = 0.0
    worst_view: int = 0
    changed_area_worst_view: int = 0
    views_evaluated: int = 0
    supersample: int = 0
    complete: bool = True
    passed: bool = True
    resource_limited: bool = False
    cancelled: bool = False

    @classmethod
    def from_c(cls, m) -> "Measurement":
        return cls(**{f.name: (bool(getattr(m, f.name)) if f.type == "bool" else getattr(m, f.name))
                      for f in fields(cls)})

    def as_tuple(self) -> tuple:
        return tuple(getattr(self, f.name) for f in fields(self))


def bounds(mesh: Mesh) -> tuple[np.ndarray, float]:
    """blitz::bounds (src/mesh.cpp:77-86) computed by the library."""
    lib = _lib.load()
    b = _lib.Bounds()
    err = _lib.error_buffer()
    desc = mesh.descriptor()
    _lib.check(lib.blitz_hip_bounds_of(C.byref(desc), C.byref(b), err, len(err)), err)
    return np.array(b.center[:], dtype=np.float32), float(b.radius)


def reduce_reuse(mesh: Mesh, target_triangles: int) -> Mesh:
    """Classical quadric reduction sharing the input vertices (host C++)."""
    lib = _lib.load()
    out = np.zeros(mesh.indices.size, np.uint32)
    count = C.c_size_t()
    err = _lib.error_buffer()
    desc = mesh.descriptor()
    _lib.check(lib.blitz_hip_reduce_reuse(C.byref(desc), int(target_triangles),
                                          out.ctypes.data_as(C.POINTER(C.c_uint32)), out.size, C.byref(count),
                                          err, len(err)), err)
    return Mesh(mesh.positions, out[: count.value].reshape(-1, 3).copy(), None, mesh.double_sided)


def reduce_reuse_device(mesh: Mesh, target_triangles: int, device: int = 0, report: dict | None = None) -> Mesh:
    """blitz::reduce with endpoint (Reuse) output on the device
    (src/simplify.cpp:43-329): the reference's triangles, in its order.
    `report`, when given, receives passes, greedy rounds and collapses."""
    lib = _lib.load()
    out = np.zeros(max(1, mesh.indices.size), np.uint32)
    count = C.c_size_t()
    err = _lib.error_buffer()
    desc = mesh.descriptor()
    rep = _lib.ReduceReportC()
    rep.struct_size = C.sizeof(rep)
    _lib.check(lib.blitz_hip_reduce_reuse_device(device, C.byref(desc), int(target_triangles),
                                                 out.ctypes.data_as(C.POINTER(C.c_uint32)), out.size,
                                                 C.byref(count), C.byref(rep), err, len(err)), err)
    if report is not None:
        report.update({name: getattr(rep, name) for name, _ in rep._fields_ if name != "struct_size"})
    return Mesh(mesh.positions, out[: count.value].reshape(-1, 3).copy(), None, mesh.double_sided)


def _bounds_c(center, radius) -> _lib.Bounds:
    b = _lib.Bounds()
    c = np.asarray(center, dtype=np.float32)
    b.center[0], b.center[1], b.center[2] = float(c[0]), float(c[1]), float(c[2])
    b.radius = float(radius)
    return b


def cameras(center, radius, screen: float, orthographic: int, perspective: int, seed: int) -> list[_lib.Camera]:
    """blitz::cameras (src/evaluate.cpp:108-124)."""
    lib = _lib.load()
    n = orthographic + perspective
    out = (_lib.Camera * max(1, n))()
    b = _bounds_c(center, radius)
    _lib.check(lib.blitz_hip_cameras(C.byref(b), screen, orthographic, perspective, seed & 0xFFFFFFFF, out, n))
    return list(out)[:n]


class DeviceMesh:
    """A mesh resident in HBM (DeviceMeshView, src/neural/device_view.hpp:16)."""

    def __init__(self, audit: "Audit", handle: C.c_void_p, mesh: Mesh):
        self.audit = audit
        self.handle = handle
        self.triangles = mesh.triangles
        self.vertices = mesh.positions.shape[0]

    def release(self) -> None:
        if self.handle:
            _lib.load().blitz_hip_mesh_release(self.handle)
            self.handle = None

    def __del__(self) -> None:
        try:
            self.release()
        except Exception:
            pass


@dataclass
class Patches:
    """Local patches of a mesh as CSR arrays (blitz_hip_local_patches):
    patch i removes faces removed[removed_offsets[i]:removed_offsets[i+1]]
    (ascending) and appends added[added_offsets[i]:added_offsets[i+1]]
    (n x 3 vertex ids)."""
    removed_offsets: np.ndarray
    removed: np.ndarray
    added_offsets: np.ndarray
    added: np.ndarray

    def __len__(self) -> int:
        return int(self.removed_offsets.size - 1)

    def subset(self, ids) -> "Patches":
        ro, rv, ao, av = [0], [], [0], []
        for i in ids:
            r = self.removed[self.removed_offsets[i]:self.removed_offsets[i + 1]]
            a = self.added[self.added_offsets[i]:self.added_offsets[i + 1]]
            rv.append(r)
            av.append(a)
            ro.append(ro[-1] + r.size)
            ao.append(ao[-1] + a.shape[0])
        return Patches(np.asarray(ro, np.uint32), np.concatenate(rv).astype(np.uint32) if rv else
                       np.zeros(0, np.uint32), np.asarray(ao, np.uint32),
                       np.concatenate(av).astype(np.uint32).reshape(-1, 3) if av else np.zeros((0, 3), np.uint32))

    def candidate(self, mesh: Mesh, i: int) -> Mesh:
        """The candidate mesh of patch i: mesh faces not in the region, in
        order, then the patch's new faces."""
        keep = np.ones(mesh.triangles, bool)
        keep[self.removed[self.removed_offsets[i]:self.removed_offsets[i + 1]]] = False
        idx = np.concatenate([mesh.indices[keep], self.added[self.added_offsets[i]:self.added_offsets[i + 1]]])
        return Mesh(mesh.positions, idx, None, mesh.double_sided)


def local_patches(mesh: Mesh, count: int, seed: int, min_faces: int = 16, max_faces: int = 200) -> Patches:
    """Seeded breadth-first face regions re-meshed by legal interior collapses."""
    lib = _lib.load()
    desc = mesh.descriptor()
    nr, na = C.c_size_t(), C.c_size_t()
    err = _lib.error_buffer()
    _lib.check(lib.blitz_hip_local_patches(C.byref(desc), count, seed, min_faces, max_faces, None, None, 0, None,
                                           None, 0, C.byref(nr), C.byref(na), err, len(err)), err)
    ro = np.zeros(count + 1, np.uint32)
    ao = np.zeros(count + 1, np.uint32)
    rv = np.zeros(max(1, nr.value), np.uint32)
    av = np.zeros((max(1, na.value), 3), np.uint32)
    u32 = C.POINTER(C.c_uint32)
    _lib.check(lib.blitz_hip_local_patches(C.byref(desc), count, seed, min_faces, max_faces, ro.ctypes.data_as(u32),
                                           rv.ctypes.data_as(u32), rv.size, ao.ctypes.data_as(u32),
                                           av.ctypes.data_as(u32), av.shape[0], C.byref(nr), C.byref(na), err,
                                           len(err)), err)
    return Patches(ro, rv[:nr.value], ao, av[:na.value])


class Binding:
    """A reference bound for patch audits (blitz_hip_audit_bind)."""

    def __init__(self, audit: "Audit", handle, mesh: Mesh, center, radius, settings: "EvalSettings"):
        self.audit, self.handle, self.mesh = audit, handle, mesh
        self.center, self.radius, self.settings = center, radius, settings

    def release(self) -> None:
        if self.handle:
            _lib.load().blitz_hip_binding_release(self.handle)
            self.handle = None


class Audit:
    """One HIP audit object (AuditCuda counterpart) on one device."""

    def __init__(self, device: int = 0, memory_mib: int = 0, cache_rasters: bool = False, face_split: int = 0,
                 autotune: bool = True):
        lib = _lib.load()
        self.lib = lib
        opts = _lib.AuditOptions()
        _lib.check(lib.blitz_hip_audit_options_init(C.byref(opts), C.sizeof(opts)))
        opts.device = device
        opts.memory_mib = memory_mib
        opts.cache_rasters = int(cache_rasters)
        opts.face_split = int(face_split)
        opts.autotune = int(autotune)
        handle = C.c_void_p()
        err = _lib.error_buffer()
        _lib.check(lib.blitz_hip_audit_create(C.byref(opts), C.byref(handle), err, len(err)), err)
        self.handle = handle
        self._meshes: list[DeviceMesh] = []
        self._bindings: list[Binding] = []

    def close(self) -> None:
        for b in self._bindings:
            b.release()
        self._bindings.clear()
        for m in self._meshes:
            m.release()
        self._meshes.clear()
        if self.handle:
            self.lib.blitz_hip_audit_destroy(self.handle)
            self.handle = None

    def __enter__(self) -> "Audit":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def upload(self, mesh: Mesh) -> DeviceMesh:
        handle = C.c_void_p()
        err = _lib.error_buffer()
        desc = mesh.descriptor()
        _lib.check(self.lib.blitz_hip_mesh_upload(self.handle, C.byref(desc), C.byref(handle), err, len(err)), err)
        d = DeviceMesh(self, handle, mesh)
        self._meshes.append(d)
        return d

    def evaluate(self, reference, candidate, center, radius, settings: EvalSettings) -> Measurement:
        """Audit of candidate against reference (host or device meshes) under
        the settings' profile: coverage, normals or attributes."""
        s, keep = settings.to_c()
        b = _bounds_c(center, radius)
        out = _lib.MeasurementC()
        err = _lib.error_buffer()
        if isinstance(reference, DeviceMesh) and isinstance(candidate, DeviceMesh):
            st = self.lib.blitz_hip_audit_evaluate_device(self.handle, reference.handle, candidate.handle,
                                                          C.byref(b), C.byref(s), C.byref(out), err, len(err))
        else:
            dr, dc = reference.descriptor(), candidate.descriptor()
            st = self.lib.blitz_hip_audit_evaluate(self.handle, C.byref(dr), C.byref(dc), C.byref(b), C.byref(s),
                                                   C.byref(out), err, len(err))
        del keep
        _lib.check(st, err)
        return Measurement.from_c(out)

    def evaluate_batch(self, reference: DeviceMesh, candidates: list[DeviceMesh], center, radius,
                       settings: EvalSettings) -> list[Measurement]:
        s, keep = settings.to_c()
        b = _bounds_c(center, radius)
        n = len(candidates)
        handles = (C.c_void_p * max(1, n))(*[c.handle for c in candidates])
        out = (_lib.MeasurementC * max(1, n))()
        err = _lib.error_buffer()
        st = self.lib.blitz_hip_audit_evaluate_batch(self.handle, reference.handle, handles, n, C.byref(b),
                                                     C.byref(s), out, err, len(err))
        del keep
        _lib.check(st, err)
        return [Measurement.from_c(out[i]) for i in range(n)]

    def bind(self, mesh: Mesh, center, radius, settings: EvalSettings) -> Binding:
        """Binds a reference for patch audits (per-view masks and tile bins in HBM)."""
        s, keep = settings.to_c()
        b = _bounds_c(center, radius)
        desc = mesh.descriptor()
        handle = C.c_void_p()
        err = _lib.error_buffer()
        st = self.lib.blitz_hip_audit_bind(self.handle, C.byref(desc), C.byref(b), C.byref(s), C.byref(handle), err,
                                           len(err))
        del keep
        _lib.check(st, err)
        binding = Binding(self, handle, mesh, center, radius, settings)
        self._bindings.append(binding)
        return binding

    def evaluate_patches(self, binding: Binding, patches: Patches, settings: EvalSettings | None = None):
        """Measurements of every patch against the bound reference, and the
        number of patches that took the full audit path."""
        s, keep = (settings or binding.settings).to_c()
        n = len(patches)
        u32 = C.POINTER(C.c_uint32)
        ro = np.ascontiguousarray(patches.removed_offsets, np.uint32)
        rv = np.ascontiguousarray(patches.removed, np.uint32)
        ao = np.ascontiguousarray(patches.added_offsets, np.uint32)
        av = np.ascontiguousarray(patches.added, np.uint32).reshape(-1)
        out = (_lib.MeasurementC * max(1, n))()
        fallbacks = C.c_uint32()
        err = _lib.error_buffer()
        st = self.lib.blitz_hip_audit_evaluate_patches(
            self.handle, binding.handle, n, ro.ctypes.data_as(u32), rv.ctypes.data_as(u32) if rv.size else None,
            ao.ctypes.data_as(u32), av.ctypes.data_as(u32) if av.size else None, C.byref(s), out,
            C.byref(fallbacks), err, len(err))
        del keep
        _lib.check(st, err)
        return [Measurement.from_c(out[i]) for i in range(n)], int(fallbacks.value)

    def raster_coverage(self, mesh: DeviceMesh, center, radius, camera: _lib.Camera, screen: float, ss: int,
                        force_two_sided: bool = False):
        """Row-major packed mask words, raster side and clipped flag; None when
        the raster exceeds the 64M-sample cap."""
        size = math.ceil(screen + 8) * ss
        capacity = (size * size + 63) // 64
        words = np.zeros(max(1, capacity), dtype=np.uint64)
        out_size = C.c_uint32()
        clipped = C.c_uint8()
        b = _bounds_c(center, radius)
        err = _lib.error_buffer()
        st = self.lib.blitz_hip_raster_coverage(self.handle, mesh.handle, C.byref(b), C.byref(camera), screen, ss,
                                                int(force_two_sided),
                                                words.ctypes.data_as(C.POINTER(C.c_uint64)), capacity,
                                                C.byref(out_size), C.byref(clipped), err, len(err))
        if st == 5:
            return None
        _lib.check(st, err)
        return words[:capacity], int(out_size.value), bool(clipped.value)

    def raster_pixels(self, mesh: DeviceMesh, center, radius, camera: _lib.Camera, screen: float, ss: int,
                      force_two_sided: bool = False):
        """Attribute raster of one view (the reference's Pixel raster),
        row-major: dict of covered, visible, depth, normal (n, 3), color
        (n, 4), material; raster side; clipped flag. None over the 64M-sample
        cap."""
        size = math.ceil(screen + 8) * ss
        n = size * size
        out = {"covered": np.zeros(n, np.uint8), "visible": np.zeros(n, np.uint8), "depth": np.zeros(n, np.float32),
               "normal": np.zeros((n, 3), np.float32), "color": np.zeros((n, 4), np.float32),
               "material": np.zeros(n, np.uint16)}
        out_size = C.c_uint32()
        clipped = C.c_uint8()
        b = _bounds_c(center, radius)
        err = _lib.error_buffer()

        def ptr(a, t):
            return a.ctypes.data_as(C.POINTER(t))
        st = self.lib.blitz_hip_raster_pixels(
            self.handle, mesh.handle, C.byref(b), C.byref(camera), screen, ss, int(force_two_sided),
            ptr(out["covered"], C.c_uint8), ptr(out["visible"], C.c_uint8), ptr(out["depth"], C.c_float),
            ptr(out["normal"], C.c_float), ptr(out["color"], C.c_float), ptr(out["material"], C.c_uint16), n,
            C.byref(out_size), C.byref(clipped), err, len(err))
        if st == 5:
            return None
        _lib.check(st, err)
        return out, int(out_size.value), bool(clipped.value)

    def overlap(self, mesh, center, radius, screen: float, orthographic: int, perspective: int = 0,
                seed: int = 0xB1172026) -> float:
        """overlap_cuda (include/blitz/neural.hpp:150-151): mean overdraw samples
        per covered pixel under the top-left fill rule; inf when any view is
        clipped. `mesh` is a Mesh or a DeviceMesh."""
        dev = mesh if isinstance(mesh, DeviceMesh) else self.upload(mesh)
        try:
            out = C.c_double()
            b = _bounds_c(center, radius)
            err = _lib.error_buffer()
            _lib.check(self.lib.blitz_hip_overlap(self.handle, dev.handle, C.byref(b), screen, orthographic,
                                                  perspective, seed & 0xFFFFFFFF, C.byref(out), err, len(err)), err)
            return float(out.value)
        finally:
            if dev is not mesh:
                dev.release()

    def coverage_distance(self, a: np.ndarray, b: np.ndarray, size: int, ss: int) -> float:
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = np.ascontiguousarray(b, dtype=np.uint64)
        out = C.c_double()
        err = _lib.error_buffer()
        _lib.check(self.lib.blitz_hip_coverage_distance(self.handle, a.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                         b.ctypes.data_as(C.POINTER(C.c_uint64)), size, ss,
                                                         C.byref(out), err, len(err)), err)
        return float(out.value)

    def last_timing(self) -> dict:
        t = _lib.AuditTiming()
        t.struct_size = C.sizeof(t)
        _lib.check(self.lib.blitz_hip_audit_last_timing(self.handle, C.byref(t)))
        return {name: getattr(t, name) for name, _ in t._fields_ if name != "struct_size"}


__all__ = ["Mesh", "EvalSettings", "Measurement", "Audit", "DeviceMesh", "bounds", "cameras", "reduce_reuse",
           "BlitzError", "Binding", "Patches", "local_patches",
           "PROFILE_COVERAGE", "PROFILE_NORMALS", "PROFILE_ATTRIBUTES"]
