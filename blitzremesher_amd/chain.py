"""Python view of the HIP LOD chain (include/blitz_hip.h, blitz_hip_generate):
the counterpart of blitz_generate_neural (reference include/blitz/blitz.h:165-167)
with explicit reuse output and proposals from the source. Every proposal is
made by the device collapse executor and every gate, offer and confirmation
is a HIP audit; nothing here computes.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .audit import Audit, Measurement, Mesh


@dataclass
class ChainSettings:
    """blitz::Settings defaults (include/blitz/remesher.hpp:32-54) for the
    fields neural generation reads. Zero base_pixels / max_lod0_delta_px
    means unset, as in blitz_settings."""
    levels: int = 8
    profile: int = 1
    objective: int = 0
    triangle_overhead_bps: int = 0
    pixels_per_meter: float = 512.0
    meters_per_unit: float = 1.0
    base_pixels: float = 0.0
    last_pixels: float = 16.0
    max_lod0_delta_px: float = 0.0
    transition: tuple = ((0.0, 2.0), (1.0, 3.0))
    normal_importance: tuple = ((0.0, 1.0), (1.0, 1.0))
    attribute_importance: tuple = ((0.0, 1.0), (1.0, 1.0))
    normal_weight: float = 180.0 / (10 * math.pi)
    color_weight: float = 4.0
    material_weight: float = 4.0
    search_orthographic: int = 42
    search_perspective: int = 12
    search_seed: int = 0xB1172024
    audit_orthographic: int = 642
    audit_perspective: int = 64
    audit_seed: int = 0xB1172026
    search_supersample: int = 4
    audit_supersample: int = 8
    max_supersample: int = 32
    max_changed_area: float = 1.0
    candidate_budget: int = 64
    beam_width: int = 8

    def to_c(self):
        """The ABI-5 blitz_settings and the arrays it borrows."""
        c = _lib.BlitzSettings()
        c.struct_size = C.sizeof(_lib.BlitzSettings)
        c.abi_version = _lib.BLITZ_ABI_VERSION
        for name in ("levels", "profile", "objective", "triangle_overhead_bps", "pixels_per_meter",
                     "meters_per_unit", "base_pixels", "last_pixels", "max_lod0_delta_px", "normal_weight",
                     "color_weight", "material_weight", "search_supersample", "audit_supersample",
                     "max_supersample", "max_changed_area", "candidate_budget", "beam_width", "search_seed",
                     "audit_seed"):
            setattr(c, name, getattr(self, name))
        c.search_ortho = self.search_orthographic
        c.search_perspective = self.search_perspective
        c.audit_ortho = self.audit_orthographic
        c.audit_perspective = self.audit_perspective
        c.max_added_vertex_bytes_bps = 0xFFFFFFFF
        c.prune = 1
        c.coupled_wedges = 1
        keep = []
        for name in ("transition", "normal_importance", "attribute_importance"):
            pts = (_lib.CurvePoint * len(getattr(self, name)))(*[_lib.CurvePoint(x, y) for x, y in getattr(self, name)])
            keep.append(pts)
            setattr(c, name, C.cast(pts, C.POINTER(_lib.CurvePoint)))
            setattr(c, name + "_count", len(pts))
        return c, keep


@dataclass
class ChainOptions:
    """blitz_neural_options_v2 fields the HIP chain reads."""
    action_trials: int = 64
    action_batch: int = 32
    ranking: int = 4  # 0 learned, 1 constant, 3 shortest edge, 4 current plane
    overdraw_tiebreak: bool = True
    origin: int = 0
    output: int = 1
    confirmation: int = 1

    def to_c(self) -> _lib.ChainOptionsC:
        o = _lib.ChainOptionsC()
        _lib.check(_lib.load().blitz_hip_chain_options_init(C.byref(o), C.sizeof(o)))
        for name in ("action_trials", "action_batch", "ranking", "origin", "output", "confirmation"):
            setattr(o, name, getattr(self, name))
        o.overdraw_tiebreak = int(self.overdraw_tiebreak)
        return o


@dataclass
class ChainLod:
    indices: np.ndarray  # (m, 3) into the source's vertices
    materials: np.ndarray | None
    screen_pixels: float
    transition_limit: float
    source_limit: float
    adjacent: Measurement
    source: Measurement

    @property
    def triangles(self) -> int:
        return int(self.indices.shape[0])


@dataclass
class ChainResult:
    status: int
    lods: list[ChainLod]
    info: dict
    finalists: list[list[int]]
    proposals: list[dict] = field(default_factory=list)


def schedule(settings: ChainSettings, center, radius: float) -> list[list[float]]:
    """blitz::schedule (host): per level [pixels, transition, source]."""
    lib = _lib.load()
    cs, keep = settings.to_c()
    b = _lib.Bounds((C.c_float * 3)(*[float(x) for x in center]), float(radius))
    out = np.zeros(3 * settings.levels, np.float64)
    err = _lib.error_buffer()
    _lib.check(lib.blitz_hip_schedule(C.byref(cs), C.byref(b), out.ctypes.data_as(C.POINTER(C.c_double)), out.size,
                                      err, len(err)), err)
    del keep
    return out.reshape(-1, 3).tolist()


def generate(audit: Audit, mesh: Mesh, settings: ChainSettings, options: ChainOptions | None = None,
             scorer=None, uv: np.ndarray | None = None) -> ChainResult:
    """blitz_hip_generate; returns the whole chain (BLITZ_BUDGET_LIMITED and
    BLITZ_CANCELLED come back as `status` 1 and 2 with a valid chain)."""
    lib = _lib.load()
    options = options or ChainOptions()
    desc = mesh.descriptor()
    uv_arr = None
    if uv is not None:
        uv_arr = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        desc.uv = _lib.BlitzStream(uv_arr.ctypes.data, uv_arr.shape[0], 8)
    cs, keep = settings.to_c()
    co = options.to_c()
    handle = C.c_void_p()
    err = _lib.error_buffer()
    rc = lib.blitz_hip_generate(audit.handle, C.byref(desc), C.byref(cs), C.byref(co),
                                scorer.handle if scorer is not None else None, C.byref(handle), err, len(err))
    if rc not in (0, 4, 5):
        _lib.check(rc, err)
    try:
        info = _lib.ChainInfoC()
        info.struct_size = C.sizeof(info)
        _lib.check(lib.blitz_hip_chain_info_of(handle, C.byref(info)))
        d = {name: getattr(info, name) for name, _ in _lib.ChainInfoC._fields_ if name != "struct_size"}
        for k in ("audit_evaluations", "rejected_gates", "area_rejected_gates"):
            d[k] = list(d[k])
        lods = []
        for i in range(info.levels):
            l = _lib.ChainLodC()
            l.struct_size = C.sizeof(l)
            _lib.check(lib.blitz_hip_chain_lod_of(handle, i, C.byref(l)))
            idx = np.ctypeslib.as_array(l.indices, (l.index_count,)).copy().reshape(-1, 3) if l.index_count else \
                np.zeros((0, 3), np.uint32)
            mats = np.ctypeslib.as_array(l.materials, (l.material_count,)).copy() if l.material_count else None
            lods.append(ChainLod(idx, mats, l.screen_pixels, l.transition_limit, l.source_limit,
                                 Measurement.from_c(l.adjacent), Measurement.from_c(l.source)))
        finalists = []
        for i in range(info.finalists):
            t = np.zeros(info.levels, np.uint32)
            _lib.check(lib.blitz_hip_chain_finalist(handle, i, t.ctypes.data_as(C.POINTER(C.c_uint32)), t.size))
            finalists.append([int(x) for x in t])
        proposals = []
        for i in range(info.proposals):
            p = _lib.ChainProposalC()
            p.struct_size = C.sizeof(p)
            _lib.check(lib.blitz_hip_chain_proposal_of(handle, i, C.byref(p)))
            proposals.append({name: getattr(p, name) for name, _ in _lib.ChainProposalC._fields_
                              if name not in ("struct_size", "reserved")})
        return ChainResult(int(info.status), lods, d, finalists, proposals)
    finally:
        lib.blitz_hip_chain_destroy(handle)
        del keep
