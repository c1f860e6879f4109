"""The library's own device sort, scan and run-length primitives (the action
state's weld, ranking and trial compaction, the scorer ranking, the patch
bins and the training gathers run on them) against numpy: stable ascending
sorts of 32- and 64-bit keys over partial bit ranges (equal keys keep input
order: numpy's stable argsort), the descending float sort with ties, -0/+0
and infinities, exclusive sums (32- and 64-bit), inclusive max and run-length
encoding, at sizes from empty and one element to several tiles and a
non-multiple of the tile."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

SIZES = [0, 1, 7, 2048, 2049, 100_003]


def primitive(kind, keys, values=None, begin_bit=0, end_bit=32, want_values=True):
    from blitzremesher_amd import _lib
    lib = _lib.load()
    keys = np.ascontiguousarray(keys)
    n = len(keys)
    out = np.zeros_like(keys)
    vout = np.zeros(n, np.uint32) if want_values else None
    runs = np.zeros(1, np.uint32)
    vin = None if values is None else np.ascontiguousarray(values, np.uint32)
    err = _lib.error_buffer()
    rc = lib.blitz_hip_device_primitive(0, kind, keys.ctypes.data if n else None,
                                        vin.ctypes.data if vin is not None and n else None, n, begin_bit, end_bit,
                                        out.ctypes.data if n else None,
                                        vout.ctypes.data if vout is not None and n else None,
                                        runs.ctypes.data, err, len(err))
    _lib.check(rc, err)
    return out, vout, int(runs[0])


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES)
def test_gpu_sort_u32_is_stable(n):
    rng = np.random.default_rng(n + 1)
    keys = rng.integers(0, 1 << 12, n, dtype=np.uint32) * np.uint32(0x10001)  # many ties
    values = rng.integers(0, 1 << 31, n, dtype=np.uint32)
    for lo, hi in ((0, 32), (0, 16), (8, 24)):
        part = (keys >> np.uint32(lo)) & np.uint32((1 << (hi - lo)) - 1) if hi - lo < 32 else keys
        order = np.argsort(part, kind="stable")
        got_k, got_v, _ = primitive(0, keys, values, lo, hi)
        assert np.array_equal(got_k, keys[order]), (lo, hi)
        assert np.array_equal(got_v, values[order]), (lo, hi)
    # Without input values the sort carries the input positions.
    _, pos, _ = primitive(0, keys, None, 0, 32)
    assert np.array_equal(pos, np.argsort(keys, kind="stable").astype(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES)
def test_gpu_sort_u64_is_stable(n):
    rng = np.random.default_rng(n + 2)
    keys = (rng.integers(0, 1 << 20, n, dtype=np.uint64) << np.uint64(40)) | rng.integers(0, 4, n, dtype=np.uint64)
    order = np.argsort(keys, kind="stable")
    got_k, got_v, _ = primitive(1, keys, None, 0, 64)
    assert np.array_equal(got_k, keys[order])
    assert np.array_equal(got_v, order.astype(np.uint32))
    keys_only, _, _ = primitive(1, keys, None, 0, 64, want_values=False)
    assert np.array_equal(keys_only, keys[order])


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES)
def test_gpu_sort_float_descending_is_stable(n):
    rng = np.random.default_rng(n + 3)
    pool = np.array([0.0, -0.0, 1.5, -1.5, np.inf, -np.inf, 3e-39, -3e-39, 7.0], np.float32)
    keys = np.where(rng.random(n) < 0.5, rng.choice(pool, n), rng.normal(0, 10, n)).astype(np.float32)
    values = np.arange(n, dtype=np.uint32)
    # Descending and stable; -0 and +0 compare equal (as the float values do).
    order = np.argsort(-keys.astype(np.float64), kind="stable")
    got_k, got_v, _ = primitive(2, keys, values)
    assert np.array_equal(got_v, values[order])
    assert np.array_equal(got_k, keys[order])


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES)
def test_gpu_scans(n):
    rng = np.random.default_rng(n + 4)
    a = rng.integers(0, 1000, n, dtype=np.uint32)
    got, _, _ = primitive(3, a, want_values=False)
    assert np.array_equal(got, (np.cumsum(a, dtype=np.uint64) - a).astype(np.uint32))
    got, _, _ = primitive(4, a, want_values=False)
    assert np.array_equal(got, np.maximum.accumulate(a) if n else a)
    b = rng.integers(0, 1 << 40, n, dtype=np.uint64)
    got, _, _ = primitive(6, b, want_values=False)
    assert np.array_equal(got, np.cumsum(b, dtype=np.uint64) - b)


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES)
def test_gpu_run_length_encode(n):
    rng = np.random.default_rng(n + 5)
    keys = np.sort(rng.integers(0, max(1, n // 3), n, dtype=np.uint64) << np.uint64(33))
    unique, counts, runs = primitive(5, keys)
    want_u, want_c = np.unique(keys, return_counts=True)
    assert runs == len(want_u)
    assert np.array_equal(unique[:runs], want_u)
    assert np.array_equal(counts[:runs], want_c.astype(np.uint32))
