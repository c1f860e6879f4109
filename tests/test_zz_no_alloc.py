"""blitz_hip_coverage_distance_noalloc (the entry point that allocates
nothing) against blitz_hip_coverage_distance on the same masks: two calls in a
row, the second on a smaller raster, plus identical, one-empty and dense-change
cases and the size limit. Device buffers for the masks and the output are
allocated here, by the test, through the HIP runtime the library links.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from blitzremesher_amd import _lib


def _hip():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


H2D, D2H = 1, 2


class DeviceArray:
    def __init__(self, hip, nbytes: int):
        self.hip = hip
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), max(8, nbytes)) == 0
        self.nbytes = nbytes

    def upload(self, a: np.ndarray) -> None:
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def download(self, a: np.ndarray) -> None:
        assert self.hip.hipMemcpy(a.ctypes.data, self.ptr, a.nbytes, D2H) == 0

    def free(self) -> None:
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def words_of(mask: np.ndarray) -> np.ndarray:
    flat = np.ascontiguousarray(mask, np.uint8).reshape(-1)
    pad = (-flat.size) % 64
    bits = np.concatenate([flat, np.zeros(pad, np.uint8)]).reshape(-1, 64)
    return (bits.astype(np.uint64) << np.arange(64, dtype=np.uint64)).sum(axis=1, dtype=np.uint64)


def discs(size: int, rng, count: int, scale: float) -> np.ndarray:
    y, x = np.mgrid[0:size, 0:size]
    m = np.zeros((size, size), bool)
    for _ in range(count):
        cx, cy, r = rng.uniform(0, size), rng.uniform(0, size), rng.uniform(1, size * scale)
        m |= (x - cx) ** 2 + (y - cy) ** 2 <= r * r
    return m


def noalloc(lib, da, db, dout, size, ss) -> float:
    err = _lib.error_buffer()
    _lib.check(lib.blitz_hip_coverage_distance_noalloc(0, da.ptr, db.ptr, size, ss, dout.ptr, None, err, len(err)),
               err)
    out = np.zeros(1, np.float64)
    dout.download(out)  # hipMemcpy waits for the queued work
    return float(out[0])


@pytest.mark.gpu
def test_noalloc_distance_matches_coverage_distance():
    from blitzremesher_amd import Audit
    lib = _lib.load()
    hip = _hip()
    rng = np.random.default_rng(2024)
    cap = (1024 * 1024 + 63) // 64 * 8
    da, db, dout = DeviceArray(hip, cap), DeviceArray(hip, cap), DeviceArray(hip, 8)
    try:
        with Audit(0) as audit:
            # Two calls in a row, the second smaller; then the special cases.
            cases = []
            for size, ss in ((1024, 8), (301, 4)):
                a = discs(size, rng, 6, 0.2)
                b = a.copy()
                b[rng.random(b.shape) < 0.002] ^= True  # scattered flips: a sparse change
                b |= discs(size, rng, 2, 0.05)
                cases.append((size, ss, a, b))
            size = 640
            a = discs(size, rng, 5, 0.2)
            cases.append((size, 8, a, a.copy()))                          # identical: 0
            cases.append((size, 8, a, np.zeros_like(a)))                  # one empty: inf
            cases.append((size, 2, a, discs(size, rng, 7, 0.2)))          # dense change
            cases.append((size, 1, np.zeros_like(a), np.zeros_like(a)))   # both empty: 0
            cases.append((9, 1, rng.random((9, 9)) < 0.5, rng.random((9, 9)) < 0.5))
            for size, ss, a, b in cases:
                wa, wb = words_of(a), words_of(b)
                da.upload(wa)
                db.upload(wb)
                got = noalloc(lib, da, db, dout, size, ss)
                want = audit.coverage_distance(wa, wb, size, ss)
                assert got == want or (math.isinf(got) and math.isinf(want)), (size, ss, got, want)
                print(size, ss, got, flush=True)
        err = _lib.error_buffer()
        assert lib.blitz_hip_coverage_distance_noalloc(0, da.ptr, db.ptr, 8193, 1, dout.ptr, None, err, len(err)) == 1
        assert lib.blitz_hip_coverage_distance_noalloc(0, da.ptr, db.ptr, 64, 0, dout.ptr, None, err, len(err)) == 1
        assert lib.blitz_hip_coverage_distance_noalloc(0, None, db.ptr, 64, 1, dout.ptr, None, err, len(err)) == 1
    finally:
        hip.hipDeviceSynchronize()
        for d in (da, db, dout):
            d.free()
