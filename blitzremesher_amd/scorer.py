"""Python view of the HIP edge-graph proposal scorer (include/blitz_hip.h).

``Scorer`` holds float32 weights (row-major [out, in] then bias per layer,
the reference layout of src/neural/internal.hpp:42-43); ``Scorer.graph``
builds one mesh's neighbour graph and vertex features on the device;
``forward`` scores every directed contraction in FP32 (deterministic) or BF16
(matrix cores), ``rank`` orders them and ``decode`` returns placements.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .audit import Mesh, _bounds_c
from .meshes import normal

FEATURES, HIDDEN, OUTPUTS = 32, 128, 4
SHAPES = [(HIDDEN, 2 * FEATURES), (HIDDEN, 2 * HIDDEN), (HIDDEN, 2 * HIDDEN), (HIDDEN, 2 * HIDDEN),
          (OUTPUTS, HIDDEN)]
FP32, BF16 = 0, 1
LEARNED, CONSTANT, SHUFFLED, SHORTEST_EDGE, CURRENT_PLANE = 0, 1, 2, 3, 4


def weight_count() -> int:
    return int(_lib.load().blitz_hip_scorer_weight_count())


def he_weights(seed: int) -> np.ndarray:
    """He-normal N(0, 2 / fan_in) matrices and zero biases, seeded."""
    parts = []
    for k, (o, i) in enumerate(SHAPES):
        parts.append((normal(seed, 1000 + k, o * i) * np.sqrt(2.0 / i)).astype(np.float32))
        parts.append(np.zeros(o, np.float32))
    return np.concatenate(parts)


class Graph:
    def __init__(self, scorer: "Scorer", handle, vertices: int, actions: int):
        self.scorer, self.handle, self.vertices, self.actions = scorer, handle, vertices, actions

    def features(self) -> np.ndarray:
        out = np.zeros((self.vertices, FEATURES), np.float32)
        _lib.check(self.scorer.lib.blitz_hip_graph_features(self.scorer.handle, self.handle,
                                                             out.ctypes.data_as(C.POINTER(C.c_float)), out.size))
        return out

    def actions_array(self) -> np.ndarray:
        src = np.zeros(self.actions, np.uint32)
        dst = np.zeros(self.actions, np.uint32)
        _lib.check(self.scorer.lib.blitz_hip_graph_actions(self.handle, src.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                           dst.ctypes.data_as(C.POINTER(C.c_uint32)), self.actions))
        return np.stack([src, dst], 1)

    def release(self) -> None:
        if self.handle:
            self.scorer.lib.blitz_hip_graph_destroy(self.handle)
            self.handle = None


class Scorer:
    def __init__(self, weights: np.ndarray, device: int = 0):
        self.lib = _lib.load()
        w = np.ascontiguousarray(weights, np.float32)
        handle = C.c_void_p()
        err = _lib.error_buffer()
        _lib.check(self.lib.blitz_hip_scorer_create(device, w.ctypes.data_as(C.POINTER(C.c_float)), w.size,
                                                    C.byref(handle), err, len(err)), err)
        self.handle = handle
        self._graphs: list[Graph] = []

    def close(self) -> None:
        for g in self._graphs:
            g.release()
        self._graphs.clear()
        if self.handle:
            self.lib.blitz_hip_scorer_destroy(self.handle)
            self.handle = None

    def __enter__(self) -> "Scorer":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def set_weights(self, weights: np.ndarray) -> None:
        w = np.ascontiguousarray(weights, np.float32)
        err = _lib.error_buffer()
        _lib.check(self.lib.blitz_hip_scorer_set_weights(self.handle, w.ctypes.data_as(C.POINTER(C.c_float)), w.size,
                                                         err, len(err)), err)

    def graph(self, mesh: Mesh, center, radius) -> Graph:
        desc = mesh.descriptor()
        b = _bounds_c(center, radius)
        handle = C.c_void_p()
        err = _lib.error_buffer()
        _lib.check(self.lib.blitz_hip_graph_create(self.handle, C.byref(desc), C.byref(b), C.byref(handle), err,
                                                   len(err)), err)
        nv, na = C.c_uint32(), C.c_uint32()
        _lib.check(self.lib.blitz_hip_graph_info(handle, C.byref(nv), C.byref(na)))
        g = Graph(self, handle, nv.value, na.value)
        self._graphs.append(g)
        return g

    def concat(self, graphs: list[Graph]) -> Graph:
        """Disjoint union of graphs of this scorer (blitz_hip_graph_concat):
        part k's vertices and actions follow those of parts 0..k-1."""
        parts = (C.c_void_p * len(graphs))(*[g.handle for g in graphs])
        handle = C.c_void_p()
        err = _lib.error_buffer()
        _lib.check(self.lib.blitz_hip_graph_concat(self.handle, parts, len(graphs), C.byref(handle), err, len(err)),
                   err)
        nv, na = C.c_uint32(), C.c_uint32()
        _lib.check(self.lib.blitz_hip_graph_info(handle, C.byref(nv), C.byref(na)))
        g = Graph(self, handle, nv.value, na.value)
        self._graphs.append(g)
        return g

    def forward(self, graph: Graph, precision: int = FP32, download: bool = True) -> np.ndarray | None:
        out = np.zeros((graph.actions, OUTPUTS), np.float32) if download else None
        err = _lib.error_buffer()
        ptr = out.ctypes.data_as(C.POINTER(C.c_float)) if download else None
        _lib.check(self.lib.blitz_hip_scorer_forward(self.handle, graph.handle, precision, ptr,
                                                     out.size if download else 0, err, len(err)), err)
        return out

    def rank(self, graph: Graph, ranking: int = LEARNED) -> np.ndarray:
        order = np.zeros(graph.actions, np.uint32)
        err = _lib.error_buffer()
        _lib.check(self.lib.blitz_hip_scorer_rank(self.handle, graph.handle, ranking,
                                                  order.ctypes.data_as(C.POINTER(C.c_uint32)), order.size, err,
                                                  len(err)), err)
        return order

    def decode(self, graph: Graph, actions: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(actions, np.uint32)
        out = np.zeros((a.size, 3), np.float32)
        err = _lib.error_buffer()
        _lib.check(self.lib.blitz_hip_scorer_decode(self.handle, graph.handle, a.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                    a.size, out.ctypes.data_as(C.POINTER(C.c_float)), err, len(err)),
                   err)
        return out

    def teacher(self, graph: Graph, actions: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """Ten teacher placement candidates per action and their targets."""
        a = np.ascontiguousarray(actions, np.uint32)
        pos = np.zeros((a.size, 10, 3), np.float32)
        tgt = np.zeros((a.size, 10, 3), np.float32)
        err = _lib.error_buffer()
        f32 = C.POINTER(C.c_float)
        _lib.check(self.lib.blitz_hip_scorer_teacher(self.handle, graph.handle, a.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                     a.size, pos.ctypes.data_as(f32), tgt.ctypes.data_as(f32), err,
                                                     len(err)), err)
        return pos, tgt

    def last_forward_ms(self) -> float:
        ms = C.c_double()
        _lib.check(self.lib.blitz_hip_scorer_last_forward_ms(self.handle, C.byref(ms)))
        return ms.value

    # ---- training
    def train_step(self, graph: Graph, actions: np.ndarray, labels: np.ndarray, targets: np.ndarray | None = None,
                   settings: TrainSettings | None = None) -> TrainResult:
        """One optimizer step on a labelled batch (blitz_hip_scorer_train_step)."""
        s = settings or TrainSettings()
        a = np.ascontiguousarray(actions, np.uint32)
        lab = np.ascontiguousarray(labels, np.uint8)
        if lab.shape != a.shape:
            raise ValueError("labels must match actions")
        t = None if targets is None else np.ascontiguousarray(targets, np.float32).reshape(a.size, 3)
        c = s.to_c()
        r = _lib.TrainResultC()
        err = _lib.error_buffer()
        _lib.check(self.lib.blitz_hip_scorer_train_step(
            self.handle, graph.handle, a.ctypes.data_as(C.POINTER(C.c_uint32)),
            lab.ctypes.data_as(C.POINTER(C.c_uint8)),
            None if t is None else t.ctypes.data_as(C.POINTER(C.c_float)), a.size, C.byref(c), C.byref(r), err,
            len(err)), err)
        return TrainResult(float(r.loss), float(r.grad_norm), bool(r.applied), int(r.step))

    def gradients(self) -> np.ndarray:
        out = np.zeros(weight_count(), np.float32)
        _lib.check(self.lib.blitz_hip_scorer_gradients(self.handle, out.ctypes.data_as(C.POINTER(C.c_float)),
                                                       out.size))
        return out

    def weights(self) -> np.ndarray:
        out = np.zeros(weight_count(), np.float32)
        _lib.check(self.lib.blitz_hip_scorer_weights(self.handle, out.ctypes.data_as(C.POINTER(C.c_float)), out.size))
        return out

    def reset_optimizer(self) -> None:
        _lib.check(self.lib.blitz_hip_scorer_reset_optimizer(self.handle))

    def last_step_ms(self) -> float:
        ms = C.c_double()
        _lib.check(self.lib.blitz_hip_scorer_last_step_ms(self.handle, C.byref(ms)))
        return ms.value


# Label bits of a training batch entry.
VALID, PREFERRED, PLACED = 1, 2, 4


@dataclass
class TrainSettings:
    """Loss and AdamW settings (defaults of training/update_cuda.hpp:47-53)."""
    margin: float = 1.0
    penalty: float = 1e-4
    lr: float = 1e-3
    decay: float = 1e-4
    max_norm: float = 1.0
    placement_weight: float = 1.0
    pool: int = 32
    precision: int = BF16

    def to_c(self) -> "_lib.TrainSettingsC":
        return _lib.TrainSettingsC(self.margin, self.penalty, self.lr, self.decay, self.max_norm,
                                   self.placement_weight, self.pool, self.precision)


@dataclass
class TrainResult:
    loss: float
    grad_norm: float
    applied: bool
    step: int
