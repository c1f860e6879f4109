"""Neural model files (include/blitz_hip.h, blitz_hip_model_*): the
reference's BLZNET01/02/03 weights format (src/neural/model_io.cpp:61-153),
read and written by libblitz_hip.so (host code; no device needed)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass
class Model:
    architecture: int  # 1 graph encoder, 2 action policy, 3 / 4 placement policies
    hidden_width: int
    use: int  # 0 unrestricted, 1 endpoint-reuse only
    weights: np.ndarray
    provenance: bytes
    sha256: str = ""  # of the whole file, when loaded


def load(path: str) -> Model:
    lib = _lib.load()
    handle = C.c_void_p()
    err = _lib.error_buffer()
    _lib.check(lib.blitz_hip_model_load(str(path).encode(), C.byref(handle), err, len(err)), err)
    try:
        info = _lib.ModelInfoC()
        info.struct_size = C.sizeof(info)
        _lib.check(lib.blitz_hip_model_info_of(handle, C.byref(info)))
        w = np.ctypeslib.as_array(info.weights, (info.weight_count,)).copy()
        prov = C.string_at(info.provenance, info.provenance_size) if info.provenance_size else b""
        return Model(int(info.architecture), int(info.hidden_width), int(info.use), w, prov,
                     lib.blitz_hip_model_sha256(handle).decode())
    finally:
        lib.blitz_hip_model_destroy(handle)


def save(path: str, model: Model) -> None:
    lib = _lib.load()
    w = np.ascontiguousarray(model.weights, np.float32).reshape(-1)
    prov = bytes(model.provenance)
    err = _lib.error_buffer()
    _lib.check(lib.blitz_hip_model_save(str(path).encode(), model.architecture, model.hidden_width, model.use,
                                        w.ctypes.data_as(C.POINTER(C.c_float)), w.size, prov, len(prov), err,
                                        len(err)), err)
