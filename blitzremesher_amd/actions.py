"""Python view of the device collapse-action state (include/blitz_hip.h,
blitz_hip_actions_*): the GPU counterpart of blitz::neural::ActionState
(src/neural/action.cpp) — legal directed half-edge collapses in enumeration
order, greedy independent batches (``independent``), trial index buffers
(``trial``) and per-action patches for the bound-reference audit
(``patches``). All work runs in libblitz_hip.so on a gfx950 device.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .audit import Mesh, Patches


def _u32p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class ActionSet:
    def __init__(self, mesh: Mesh, uv: np.ndarray | None = None, device: int = 0,
                 tangents: np.ndarray | None = None):
        self.lib = _lib.load()
        self.mesh = mesh
        self.uv = None if uv is None else np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        self.tangents = None if tangents is None else np.ascontiguousarray(tangents, np.float32).reshape(-1, 4)
        desc = mesh.descriptor()
        if self.uv is not None:
            desc.uv = _lib.BlitzStream(self.uv.ctypes.data, self.uv.shape[0], 8)
        if self.tangents is not None:
            desc.tangents = _lib.BlitzStream(self.tangents.ctypes.data, self.tangents.shape[0], 16)
        handle = C.c_void_p()
        err = _lib.error_buffer()
        _lib.check(self.lib.blitz_hip_actions_create(device, C.byref(desc), C.byref(handle), err, len(err)), err)
        self.handle = handle
        edges, legal, ms = C.c_uint32(), C.c_uint32(), C.c_double()
        _lib.check(self.lib.blitz_hip_actions_info(self.handle, C.byref(edges), C.byref(legal), C.byref(ms)))
        self.edges, self.build_ms = int(edges.value), float(ms.value)
        pairs = np.zeros(2 * max(1, legal.value), np.uint32)
        _lib.check(self.lib.blitz_hip_actions_legal(self.handle, _u32p(pairs), int(legal.value), err, len(err)), err)
        self.legal = pairs[: 2 * legal.value].reshape(-1, 2).copy()

    def close(self) -> None:
        if self.handle:
            self.lib.blitz_hip_actions_destroy(self.handle)
            self.handle = None

    def __enter__(self) -> "ActionSet":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rank(self, ranking: int, scorer=None, graph=None) -> np.ndarray:
        """Rank order of the legal actions (0 learned from `graph` after a
        scorer forward, 1 constant, 3 shortest edge, 4 current plane)."""
        order = np.zeros(max(1, len(self.legal)), np.uint32)
        err = _lib.error_buffer()
        _lib.check(self.lib.blitz_hip_actions_rank(self.handle, ranking, scorer.handle if scorer else None,
                                                   graph.handle if graph else None, _u32p(order), order.size, err,
                                                   len(err)), err)
        return order[: len(self.legal)].copy()

    def vertex_features(self) -> np.ndarray:
        """blitz::neural::graph's 24 channels per vertex (n, 24) of this state's mesh."""
        out = np.zeros((self.mesh.positions.shape[0], 24), np.float32)
        err = _lib.error_buffer()
        _lib.check(self.lib.blitz_hip_actions_vertex_features(self.handle, out.ctypes.data_as(C.POINTER(C.c_float)),
                                                              out.size, err, len(err)), err)
        return out

    def features(self, conditions=None, width: int = 80, input_triangles: int = 0) -> np.ndarray:
        """ActionState::actions' feature rows (legal, width) for the eight conditions."""
        c = np.zeros(8, np.float32) if conditions is None else np.ascontiguousarray(conditions, np.float32)
        out = np.zeros((max(1, len(self.legal)), width), np.float32)
        err = _lib.error_buffer()
        _lib.check(self.lib.blitz_hip_actions_features(self.handle, c.ctypes.data_as(C.POINTER(C.c_float)), width,
                                                       input_triangles, out.ctypes.data_as(C.POINTER(C.c_float)),
                                                       out.size, err, len(err)), err)
        return out[: len(self.legal)]

    def rank_policy(self, policy: "Policy", conditions=None, input_triangles: int = 0,
                    preserve_uv: bool = False) -> np.ndarray:
        """Learned order of the reference's GPU action state: the policy's
        first output on each legal action's feature row, descending."""
        c = np.zeros(8, np.float32) if conditions is None else np.ascontiguousarray(conditions, np.float32)
        order = np.zeros(max(1, len(self.legal)), np.uint32)
        err = _lib.error_buffer()
        _lib.check(self.lib.blitz_hip_actions_rank_policy(self.handle, policy.handle,
                                                          c.ctypes.data_as(C.POINTER(C.c_float)), input_triangles,
                                                          int(preserve_uv), _u32p(order), order.size, err, len(err)),
                   err)
        return order[: len(self.legal)].copy()

    def placements(self, policy: "Policy | None" = None, conditions=None, input_triangles: int = 0,
                   preserve_uv: bool = False, quantization=None, exact_bits: np.ndarray | None = None) -> np.ndarray:
        """Placed position and destination-wedge normals (legal, 9) of every legal
        action; `quantization` = (low[3], extent[3]) snaps to its unorm16 grid."""
        c = np.zeros(8, np.float32) if conditions is None else np.ascontiguousarray(conditions, np.float32)
        o = _lib.PlacementOptionsC()
        _lib.check(self.lib.blitz_hip_placement_options_init(C.byref(o), C.sizeof(o)))
        o.preserve_uv = int(preserve_uv)
        if quantization is not None:
            o.quantized = 1
            o.low[:] = [float(v) for v in quantization[0]]
            o.extent[:] = [float(v) for v in quantization[1]]
        bits = None
        if exact_bits is not None:
            bits = np.ascontiguousarray(exact_bits, np.uint32)
            o.exact_bits = _u32p(bits)
            o.exact_words = bits.size
        out = np.zeros((max(1, len(self.legal)), 9), np.float32)
        err = _lib.error_buffer()
        _lib.check(self.lib.blitz_hip_actions_placements(self.handle, policy.handle if policy else None,
                                                         c.ctypes.data_as(C.POINTER(C.c_float)), input_triangles,
                                                         C.byref(o), out.ctypes.data_as(C.POINTER(C.c_float)),
                                                         out.size, err, len(err)), err)
        return out[: len(self.legal)]

    def mapping(self) -> np.ndarray:
        """(legal, 5): wedges, from0, to0, from1, to1 (0xFFFFFFFF when absent)."""
        out = np.zeros((max(1, len(self.legal)), 5), np.uint32)
        err = _lib.error_buffer()
        _lib.check(self.lib.blitz_hip_actions_mapping(self.handle, _u32p(out), out.size, err, len(err)), err)
        return out[: len(self.legal)]

    def independent(self, order: np.ndarray, maximum: int, maximum_removed: int) -> np.ndarray:
        """ActionState::independent over legal-action indices in rank order."""
        order = np.ascontiguousarray(order, np.uint32)
        out = np.zeros(64, np.uint32)
        count = C.c_size_t()
        err = _lib.error_buffer()
        _lib.check(self.lib.blitz_hip_actions_independent(self.handle, _u32p(order), order.size, maximum,
                                                          maximum_removed, _u32p(out), C.byref(count), err, len(err)),
                   err)
        return out[: count.value].copy()

    def trial(self, rows: np.ndarray) -> np.ndarray:
        """Index buffer (m, 3) with the legal actions `rows` applied."""
        rows = np.ascontiguousarray(rows, np.uint32)
        out = np.zeros(self.mesh.indices.size, np.uint32)
        count = C.c_size_t()
        err = _lib.error_buffer()
        _lib.check(self.lib.blitz_hip_actions_trial(self.handle, _u32p(rows), rows.size, _u32p(out), out.size,
                                                    C.byref(count), err, len(err)), err)
        return out[: count.value].reshape(-1, 3).copy()

    def patches(self, rows: np.ndarray) -> Patches:
        """Each legal action as a patch of the source mesh."""
        rows = np.ascontiguousarray(rows, np.uint32)
        nr, na = C.c_size_t(), C.c_size_t()
        err = _lib.error_buffer()
        null = C.POINTER(C.c_uint32)()
        _lib.check(self.lib.blitz_hip_actions_patches(self.handle, _u32p(rows), rows.size, null, null, 0, null, null,
                                                      0, C.byref(nr), C.byref(na), err, len(err)), err)
        ro = np.zeros(rows.size + 1, np.uint32)
        ao = np.zeros(rows.size + 1, np.uint32)
        rv = np.zeros(max(1, nr.value), np.uint32)
        av = np.zeros(max(1, 3 * na.value), np.uint32)
        _lib.check(self.lib.blitz_hip_actions_patches(self.handle, _u32p(rows), rows.size, _u32p(ro), _u32p(rv),
                                                      rv.size, _u32p(ao), _u32p(av), av.size // 3, C.byref(nr),
                                                      C.byref(na), err, len(err)), err)
        return Patches(ro, rv[: nr.value].copy(), ao, av[: 3 * na.value].reshape(-1, 3).copy())


def execute_actions(audit, mesh: Mesh, center, radius, gate, target: int, budget: int, ranking: int,
                    batch: int, uv: np.ndarray | None = None, scorer=None):
    """neural::execute_actions (src/neural/action.cpp:453-525) on the device:
    every trial gated by an exact audit against `mesh` under the EvalSettings
    `gate`; ranking 0 orders by the BF16 logits of `scorer`. Returns (index
    buffer (m, 3), stats dict)."""
    from .audit import _bounds_c
    lib = _lib.load()
    desc = mesh.descriptor()
    if uv is not None:
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        desc.uv = _lib.BlitzStream(uv.ctypes.data, uv.shape[0], 8)
    s, keep = gate.to_c()
    b = _bounds_c(center, radius)
    out = np.zeros(mesh.indices.size, np.uint32)
    count = C.c_size_t()
    st = _lib.ActionStats()
    st.struct_size = C.sizeof(st)
    err = _lib.error_buffer()
    _lib.check(lib.blitz_hip_execute_actions(audit.handle, C.byref(desc), C.byref(b), C.byref(s), int(target), budget,
                                             ranking, scorer.handle if scorer is not None else None, batch,
                                             _u32p(out), out.size, C.byref(count), C.byref(st), err, len(err)), err)
    del keep
    stats = {k: getattr(st, k) for k, _ in st._fields_ if k != "struct_size"}
    return out[: count.value].reshape(-1, 3).copy(), stats


__all__ = ["ActionSet", "execute_actions"]


# Inputs and outputs of the policy architectures (src/neural/placement.hpp:17-26).
POLICY_SHAPES = {2: (80, 3), 3: (128, 12), 4: (128, 12)}


def policy_weight_count(architecture: int, hidden: int = 64) -> int:
    i, o = POLICY_SHAPES[architecture]
    return hidden * (i + 1) + hidden * (hidden + 1) + o * (hidden + 1)


class Policy:
    """ActionCuda (src/neural/action.hpp:70-87): the action (architecture 2)
    or placement (3, 4) policy on the device, three compensated FP32 layers."""

    def __init__(self, weights: np.ndarray, architecture: int = 2, hidden: int = 64, device: int = 0):
        self.lib = _lib.load()
        self.architecture, self.hidden = architecture, hidden
        self.inputs, self.outputs = POLICY_SHAPES.get(architecture, (0, 0))
        w = np.ascontiguousarray(weights, np.float32).reshape(-1)
        handle = C.c_void_p()
        err = _lib.error_buffer()
        _lib.check(self.lib.blitz_hip_policy_create(device, architecture, hidden,
                                                    w.ctypes.data_as(C.POINTER(C.c_float)), w.size,
                                                    C.byref(handle), err, len(err)), err)
        self.handle = handle

    @classmethod
    def from_model(cls, path: str, device: int = 0) -> "Policy":
        """The policy of a BLZNET action or placement model file."""
        self = cls.__new__(cls)
        self.lib = _lib.load()
        err = _lib.error_buffer()
        model = C.c_void_p()
        _lib.check(self.lib.blitz_hip_model_load(str(path).encode(), C.byref(model), err, len(err)), err)
        try:
            info = _lib.ModelInfoC()
            info.struct_size = C.sizeof(info)
            _lib.check(self.lib.blitz_hip_model_info_of(model, C.byref(info)))
            self.architecture, self.hidden = int(info.architecture), int(info.hidden_width)
            self.inputs, self.outputs = POLICY_SHAPES.get(self.architecture, (0, 0))
            handle = C.c_void_p()
            _lib.check(self.lib.blitz_hip_policy_from_model(device, model, C.byref(handle), err, len(err)), err)
            self.handle = handle
        finally:
            self.lib.blitz_hip_model_destroy(model)
        return self

    def predict(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.inputs)
        out = np.zeros((max(1, x.shape[0]), self.outputs), np.float32)
        err = _lib.error_buffer()
        _lib.check(self.lib.blitz_hip_policy_predict(self.handle, x.ctypes.data_as(C.POINTER(C.c_float)), x.shape[0],
                                                     out.ctypes.data_as(C.POINTER(C.c_float)), out.size, err,
                                                     len(err)), err)
        return out[: x.shape[0]]

    def close(self) -> None:
        if self.handle:
            self.lib.blitz_hip_policy_destroy(self.handle)
            self.handle = None

    def __enter__(self) -> "Policy":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
