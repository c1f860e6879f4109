"""Neural model files: the reference's BLZNET01/02/03 format
(src/neural/model_io.cpp:61-153) read and written by libblitz_hip.so.

CPU (host code, no device): the reference's own model file
(tests/fixtures/neural-cancellation/model.blzn of the reference, kept as
tests/golden/reference_placement_model.blzn) loads with its declared shape
and re-saves to the same bytes; a file assembled here from the documented
layout with Python's hashlib loads; the SHA-256 trailer equals hashlib's
over payload lengths crossing every padding boundary; the use-scope
contracts of the reference's tests/neural_model_scope.cpp hold; corrupt
files are rejected.
GPU: the reference model's placement policy runs on the device, bit-exact
to the compensated restatement.
"""
from __future__ import annotations

import hashlib
import struct
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent))
sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
import scorer_oracle as so  # noqa: E402

from blitzremesher_amd import model as mio  # noqa: E402
from blitzremesher_amd._lib import BlitzError  # noqa: E402
from blitzremesher_amd.actions import policy_weight_count  # noqa: E402

FIXTURE = Path(__file__).resolve().parent / "golden" / "reference_placement_model.blzn"
ENCODER_WEIGHTS = 64 * 49 + 64 * 129 + 64 * 129 + 64 * 73 + 4 * 65  # layers 48,128,128,72,64 -> 64,64,64,64,4


def blzn(magic: bytes, architecture: int, values: np.ndarray, provenance: bytes, width=None, use=None) -> bytes:
    """A model file from the documented layout (little-endian u32 fields,
    float32 weights, hex SHA-256 of everything before the trailer)."""
    head = magic + struct.pack("<III", architecture, values.size, len(provenance))
    if width is not None:
        head += struct.pack("<I", width)
    if use is not None:
        head += struct.pack("<I", use)
    payload = head + provenance + np.ascontiguousarray(values, "<f4").tobytes()
    return payload + hashlib.sha256(payload).hexdigest().encode()


def test_reference_model_file_loads():
    data = FIXTURE.read_bytes()
    m = mio.load(str(FIXTURE))
    assert (m.architecture, m.hidden_width, m.use) == (3, 64, 0)
    assert m.weights.size == policy_weight_count(3, 64) == 13196
    assert len(m.provenance) == 16 and np.isfinite(m.weights).all()
    assert m.sha256 == hashlib.sha256(data).hexdigest()
    # The weights are the file's float32 payload after the 20-byte header and provenance.
    assert np.array_equal(m.weights, np.frombuffer(data[36:36 + 4 * 13196], "<f4"))


def test_resaved_reference_model_is_byte_identical(tmp_path):
    out = tmp_path / "again.blzn"
    mio.save(str(out), mio.load(str(FIXTURE)))
    assert out.read_bytes() == FIXTURE.read_bytes()


@pytest.mark.parametrize("magic,width,use", [(b"BLZNET01", None, None), (b"BLZNET02", 128, None),
                                             (b"BLZNET03", 256, 1)])
def test_file_from_documented_layout_loads(tmp_path, magic, width, use):
    arch = 4 if use else 2
    values = np.random.default_rng(5).standard_normal(policy_weight_count(arch, width or 64)).astype(np.float32)
    path = tmp_path / "m.blzn"
    path.write_bytes(blzn(magic, arch, values, b"provenance text", width, use))
    m = mio.load(str(path))
    assert (m.architecture, m.hidden_width, m.use) == (arch, width or 64, use or 0)
    assert np.array_equal(m.weights, values) and m.provenance == b"provenance text"
    out = tmp_path / "re.blzn"
    mio.save(str(out), m)
    assert out.read_bytes() == path.read_bytes()


def test_checksum_is_sha256_over_every_padding_boundary(tmp_path):
    # Payload length = 20 + provenance + 4 * weights: provenance lengths 0..130
    # step the length through all residues mod 64 (SHA-256 pads to 64 bytes).
    values = np.linspace(-1, 1, policy_weight_count(2, 64)).astype(np.float32)
    for n in range(131):
        prov = bytes((i * 7 + n) % 251 for i in range(n))
        path = tmp_path / f"m{n}.blzn"
        mio.save(str(path), mio.Model(2, 64, 0, values, prov))
        data = path.read_bytes()
        assert data[-64:].decode() == hashlib.sha256(data[:-64]).hexdigest(), n
        assert mio.load(str(path)).sha256 == hashlib.sha256(data).hexdigest()


@pytest.mark.parametrize("width", [64, 128, 256])
def test_use_scope_contracts(tmp_path, width):
    # tests/neural_model_scope.cpp of the reference, restated.
    a, b = tmp_path / "a.blzn", tmp_path / "b.blzn"
    values = np.full(policy_weight_count(4, width), 0.013, np.float32)
    prov = b"x" * 65536 if width == 256 else b"scope fixture"
    mio.save(str(a), mio.Model(4, width, 0, values, prov))
    legacy = mio.load(str(a))
    assert legacy.use == 0 and legacy.hidden_width == width and np.array_equal(legacy.weights, values)
    mio.save(str(b), legacy)
    assert a.read_bytes() == b.read_bytes()
    legacy_size = a.stat().st_size
    mio.save(str(a), mio.Model(4, width, 1, values, prov))
    endpoint = mio.load(str(a))
    assert endpoint.use == 1 and endpoint.hidden_width == width and np.array_equal(endpoint.weights, values)
    assert a.stat().st_size == legacy_size + (8 if width == 64 else 4)
    mio.save(str(b), endpoint)
    assert a.read_bytes() == b.read_bytes()
    with pytest.raises(BlitzError):
        mio.save(str(a), mio.Model(4, width, 2, values, prov))
    with pytest.raises(BlitzError):
        mio.save(str(a), mio.Model(3, width, 1, np.full(policy_weight_count(3, width), 0.013, np.float32), prov))
    # An unknown serialized scope fails even with a correct checksum.
    data = bytearray(b.read_bytes())
    data[24:28] = struct.pack("<I", 2)
    data[-64:] = hashlib.sha256(bytes(data[:-64])).hexdigest().encode()
    a.write_bytes(bytes(data))
    with pytest.raises(BlitzError):
        mio.load(str(a))


def test_encoder_architecture_round_trips(tmp_path):
    values = np.random.default_rng(2).standard_normal(ENCODER_WEIGHTS).astype(np.float32)
    path = tmp_path / "encoder.blzn"
    mio.save(str(path), mio.Model(1, 64, 0, values, b""))
    m = mio.load(str(path))
    assert m.architecture == 1 and np.array_equal(m.weights, values)
    with pytest.raises(BlitzError):
        mio.save(str(path), mio.Model(1, 128, 0, values, b""))


def test_corrupt_files_are_rejected(tmp_path):
    good = FIXTURE.read_bytes()
    payload = good[:-64]

    def fixed(p: bytes) -> bytes:
        return p + hashlib.sha256(p).hexdigest().encode()

    nan = bytearray(payload)
    nan[36:40] = struct.pack("<f", float("nan"))
    count = bytearray(payload)
    count[12:16] = struct.pack("<I", 13195)
    cases = {
        "truncated": good[:-1],
        "too small": good[:83],
        "checksum": bytes([good[0]]) + good[1:100] + bytes([good[100] ^ 1]) + good[101:],
        "magic": fixed(b"BLZNET09" + payload[8:]),
        "count": fixed(bytes(count)),
        "nonfinite": fixed(bytes(nan)),
        "architecture": fixed(payload[:8] + struct.pack("<I", 7) + payload[12:]),
    }
    for name, data in cases.items():
        path = tmp_path / f"{name}.blzn"
        path.write_bytes(data)
        with pytest.raises(BlitzError):
            mio.load(str(path))
    with pytest.raises(BlitzError):
        mio.load(str(tmp_path / "missing.blzn"))


@pytest.mark.gpu
def test_gpu_reference_model_policy_matches_restatement():
    from blitzremesher_amd.actions import Policy
    m = mio.load(str(FIXTURE))
    rows = np.random.default_rng(9).standard_normal((777, 128)).astype(np.float32)
    with Policy.from_model(str(FIXTURE)) as p:
        assert (p.architecture, p.inputs, p.outputs) == (3, 128, 12)
        got = p.predict(rows)
    want = so.policy_forward(rows, m.weights, 64, 12)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
