"""CPU oracle of the edge-graph proposal scorer.

TEST INFRASTRUCTURE, NOT PRODUCT CODE: only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module, as the checker.

The network is the build's edge-graph scorer (the reference's production
policy is a different, independent-action MLP; see the header comment of
blitzremesher_amd/csrc/scorer_kernels.hip). Its oracle follows the structure
of the reference's FP64 numeric oracle (src/neural/numeric.cpp:52-122): every
layer is recomputed in float64 from float32 weights and inputs, with
mean-aggregated neighbours (src/neural/inference.cu:18-30). `forward_fp32`
restates, operation for operation, the library's deterministic FP32 mode,
whose linear layers follow the reference's compensated kernel
(src/neural/inference.cu:75-108): eight strided Kahan partial sums per output
combined in fixed order with the bias.
"""
from __future__ import annotations

import numpy as np

FEATURES = 32
HIDDEN = 128
LAYERS = 3
EDGE_IN = 2 * HIDDEN
OUTPUTS = 4  # logit, placement xyz

# (out, in) of every linear layer, in weight-buffer order.
SHAPES = [(HIDDEN, 2 * FEATURES), (HIDDEN, 2 * HIDDEN), (HIDDEN, 2 * HIDDEN), (HIDDEN, EDGE_IN),
          (OUTPUTS, HIDDEN)]


def weight_count() -> int:
    return sum(o * (i + 1) for o, i in SHAPES)


def split_weights(w: np.ndarray) -> list[tuple[np.ndarray, np.ndarray]]:
    """Row-major [out, in] weights then bias per layer (src/neural/internal.hpp:42-43)."""
    out, at = [], 0
    for o, i in SHAPES:
        W = w[at:at + o * i].reshape(o, i)
        at += o * i
        b = w[at:at + o]
        at += o
        out.append((W, b))
    assert at == w.size
    return out


# ---------------------------------------------------------------- topology --

def topology(indices: np.ndarray, vertices: int):
    """Undirected neighbour CSR (sorted, unique), vertex-face CSR (faces in
    index order) and the boundary-edge flags used by the features."""
    f = indices.astype(np.int64)
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    e = e[e[:, 0] != e[:, 1]]
    key = np.sort(e, axis=1)
    uniq, counts = np.unique(key, axis=0, return_counts=True)
    both = np.concatenate([uniq, uniq[:, ::-1]])
    order = np.lexsort((both[:, 1], both[:, 0]))
    both = both[order]
    offsets = np.zeros(vertices + 1, np.int64)
    np.add.at(offsets, both[:, 0] + 1, 1)
    offsets = np.cumsum(offsets)
    neighbors = both[:, 1]
    boundary_edges = uniq[counts == 1]
    vf_v = f.reshape(-1)
    vf_f = np.repeat(np.arange(len(f)), 3)
    order = np.lexsort((vf_f, vf_v))
    face_offsets = np.zeros(vertices + 1, np.int64)
    np.add.at(face_offsets, vf_v + 1, 1)
    face_offsets = np.cumsum(face_offsets)
    return offsets, neighbors, face_offsets, vf_f[order], boundary_edges


# ---------------------------------------------------------------- features --

def features(positions: np.ndarray, indices: np.ndarray, center, diameter: float) -> np.ndarray:
    """32 per-vertex features in float64 (cast to float32 by the caller);
    layout documented in blitzremesher_amd/csrc/scorer_kernels.hip."""
    n = positions.shape[0]
    offsets, neighbors, face_offsets, faces, boundary_edges = topology(indices, n)
    p = (positions.astype(np.float64) - np.asarray(center, np.float64)) / diameter
    f = indices.astype(np.int64)
    a, b, c = p[f[:, 0]], p[f[:, 1]], p[f[:, 2]]
    cr = np.cross(b - a, c - a)
    area = np.sqrt((cr * cr).sum(1))
    safe = np.where(area > 0, area, 1.0)
    nf = np.where(area[:, None] > 0, cr / safe[:, None], 0.0)
    d = -(nf * a).sum(1)
    plane = np.concatenate([nf, d[:, None]], 1)
    iu = [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]
    qf = np.stack([plane[:, i] * plane[:, j] * area for i, j in iu], 1)
    deg_f = np.diff(face_offsets)
    normal = np.zeros((n, 3))
    quad = np.zeros((n, 10))
    tot = np.zeros(n)
    # Fixed order: incident faces in ascending face index.
    for slot in range(int(deg_f.max(initial=0))):
        has = deg_f > slot
        v = np.nonzero(has)[0]
        fi = faces[face_offsets[v] + slot]
        normal[v] += cr[fi]
        quad[v] += qf[fi]
        tot[v] += area[fi]
    nl = np.sqrt((normal * normal).sum(1))
    unit = np.where(nl[:, None] > 0, normal / np.where(nl > 0, nl, 1)[:, None], 0.0)
    deg = np.diff(offsets)
    nbr_sum = np.zeros((n, 3))
    edge_sum = np.zeros(n)
    for slot in range(int(deg.max(initial=0))):
        v = np.nonzero(deg > slot)[0]
        w = neighbors[offsets[v] + slot]
        nbr_sum[v] += p[w]
        dd = p[w] - p[v]
        edge_sum[v] += np.sqrt((dd * dd).sum(1))
    degc = np.maximum(deg, 1)
    mean_edge = edge_sum / degc
    lap = nbr_sum / degc[:, None] - p
    curvature = np.clip((unit * lap).sum(1) / np.maximum(mean_edge, 1e-30), -1.0, 1.0)
    bcount = np.zeros(n)
    np.add.at(bcount, boundary_edges[:, 0], 1)
    np.add.at(bcount, boundary_edges[:, 1], 1)
    x = np.zeros((n, FEATURES))
    x[:, 0:3] = p
    x[:, 3:6] = unit
    x[:, 6:16] = quad / np.maximum(tot, 1e-30)[:, None]
    x[:, 16] = curvature
    x[:, 17] = deg / 16.0
    x[:, 18] = (bcount > 0).astype(np.float64)
    x[:, 19] = 0.0  # UV stream present
    x[:, 20] = np.minimum(mean_edge * 64.0, 4.0)
    x[:, 21] = deg_f / 8.0
    x[:, 22] = bcount / 4.0
    x[:, 23:26] = 1.0  # linear colour (white when absent)
    x[:, 26:28] = 0.0  # uv / 16
    x[:, 28] = 0.0  # normal stream present
    x[:, 29] = 0.0  # colour stream present
    x[:, 30] = np.minimum(np.sqrt(tot) / np.maximum(mean_edge, 1e-30), 4.0)
    x[:, 31] = 1.0
    return x


# ----------------------------------------------------------------- network --

def actions(offsets: np.ndarray, neighbors: np.ndarray) -> np.ndarray:
    """Directed contractions, one per CSR entry: (source, destination)."""
    src = np.repeat(np.arange(len(offsets) - 1), np.diff(offsets))
    return np.stack([src, neighbors], 1)


def forward_fp64(x: np.ndarray, offsets, neighbors, acts, weights: np.ndarray) -> np.ndarray:
    layers = [(W.astype(np.float64), b.astype(np.float64)) for W, b in split_weights(weights)]
    h = x.astype(np.float64)
    deg = np.maximum(np.diff(offsets), 1)
    src = np.repeat(np.arange(len(offsets) - 1), np.diff(offsets))
    for W, b in layers[:LAYERS]:
        agg = np.zeros_like(h)
        np.add.at(agg, src, h[neighbors])
        z = np.concatenate([h, agg / deg[:, None]], 1)
        h = np.maximum(z @ W.T + b, 0.0)
    W1, b1 = layers[3]
    W2, b2 = layers[4]
    z = np.concatenate([h[acts[:, 0]], h[acts[:, 1]]], 1)
    g = np.maximum(z @ W1.T + b1, 0.0)
    return g @ W2.T + b2


def _compensated_linear(x: np.ndarray, W: np.ndarray, b: np.ndarray, relu: bool) -> np.ndarray:
    f32 = np.float32
    rows, k_in = x.shape
    out = W.shape[0]
    sums = np.zeros((8, rows, out), f32)
    corrs = np.zeros((8, rows, out), f32)
    for lane in range(8):
        s = np.zeros((rows, out), f32)
        c = np.zeros((rows, out), f32)
        for k in range(lane, k_in, 8):
            term = (x[:, k][:, None] * W[None, :, k]).astype(f32) - c
            nxt = s + term
            c = (nxt - s) - term
            s = nxt
        sums[lane], corrs[lane] = s, c
    total = np.broadcast_to(b[None, :], (rows, out)).astype(f32).copy()
    err = np.zeros((rows, out), f32)
    for j in range(8):
        term = sums[j] - err
        nxt = total + term
        err = (nxt - total) - term
        total = nxt
        term = -corrs[j] - err
        nxt = total + term
        err = (nxt - total) - term
        total = nxt
    if relu:
        total = np.where(total < 0, f32(0), total)
    return total


def policy_forward(x: np.ndarray, weights: np.ndarray, hidden: int, outputs: int) -> np.ndarray:
    """ActionCuda::predict (src/neural/inference.cu:236-290): three
    compensated layers in[=x.shape[1]] -> hidden -> hidden -> outputs, ReLU
    after the first two, weights [out,in] then bias per layer."""
    x = np.ascontiguousarray(x, np.float32)
    w = np.ascontiguousarray(weights, np.float32)
    at = 0
    for layer, (k_in, k_out) in enumerate(((x.shape[1], hidden), (hidden, hidden), (hidden, outputs))):
        W = w[at:at + k_in * k_out].reshape(k_out, k_in)
        b = w[at + k_in * k_out:at + (k_in + 1) * k_out]
        at += (k_in + 1) * k_out
        x = _compensated_linear(x, W, b, layer < 2)
    assert at == w.size
    return x


def forward_fp32(x: np.ndarray, offsets, neighbors, acts, weights: np.ndarray) -> np.ndarray:
    """Bit-exact restatement of the library's deterministic FP32 mode."""
    f32 = np.float32
    layers = split_weights(weights.astype(f32))
    h = x.astype(f32)
    deg = np.diff(offsets)
    n = len(deg)
    for W, b in layers[:LAYERS]:
        agg = np.zeros_like(h)
        for slot in range(int(deg.max(initial=0))):
            v = np.nonzero(deg > slot)[0]
            agg[v] = agg[v] + h[neighbors[offsets[v] + slot]]
        agg = agg / np.maximum(deg, 1).astype(f32)[:, None]
        z = np.concatenate([h, agg], 1)
        h = _compensated_linear(z, W, b, True)
        assert h.shape[0] == n
    W1, b1 = layers[3]
    W2, b2 = layers[4]
    z = np.concatenate([h[acts[:, 0]], h[acts[:, 1]]], 1)
    g = _compensated_linear(z, W1, b1, True)
    return _compensated_linear(g, W2, b2, False)


def receptive_field(offsets, neighbors, vertices: np.ndarray) -> list[np.ndarray]:
    """Vertex sets needed by a forward pass that only has to produce the
    final embeddings of `vertices`: entry L holds the vertices whose layer
    input is needed L hops out (entry 0 = `vertices`)."""
    sets = [np.unique(vertices)]
    for _ in range(LAYERS):
        v = sets[-1]
        nb = np.concatenate([neighbors[offsets[i]:offsets[i + 1]] for i in v]) if len(v) else v
        sets.append(np.union1d(v, nb))
    return sets


def forward_fp32_subset(x: np.ndarray, offsets, neighbors, acts, weights: np.ndarray) -> np.ndarray:
    """forward_fp32 for the actions `acts` only: each message-passing layer is
    evaluated on the rows its successors need (the receptive field of the
    actions' endpoints). Every row's arithmetic is the one forward_fp32
    performs, so the outputs are the same bits."""
    f32 = np.float32
    layers = split_weights(weights.astype(f32))
    sets = receptive_field(offsets, neighbors, acts.reshape(-1))
    deg = np.diff(offsets)
    h = x.astype(f32).copy()
    for i, (W, b) in enumerate(layers[:LAYERS]):
        rows = sets[LAYERS - 1 - i]
        agg = np.zeros((len(rows), h.shape[1]), f32)
        d = deg[rows]
        for slot in range(int(d.max(initial=0))):
            sel = np.nonzero(d > slot)[0]
            agg[sel] = agg[sel] + h[neighbors[offsets[rows[sel]] + slot]]
        agg = agg / np.maximum(d, 1).astype(f32)[:, None]
        z = np.concatenate([h[rows], agg], 1)
        out = _compensated_linear(z, W, b, True)
        h = np.zeros((len(deg), out.shape[1]), f32)
        h[rows] = out
    W1, b1 = layers[3]
    W2, b2 = layers[4]
    z = np.concatenate([h[acts[:, 0]], h[acts[:, 1]]], 1)
    g = _compensated_linear(z, W1, b1, True)
    return _compensated_linear(g, W2, b2, False)


def rank(logits: np.ndarray) -> np.ndarray:
    """Descending score order, ties by ascending action index."""
    return np.lexsort((np.arange(len(logits)), -logits.astype(np.float64)))


# Vec3 helpers of src/neural/action_gpu.cu:16-38: float components; mul3
# scales in double and rounds to float; dot3 and len3 are double.
_f32 = np.float32


def _add3(a, b):
    return (a.astype(_f32) + b.astype(_f32)).astype(_f32)


def _sub3(a, b):
    return (a.astype(_f32) - b.astype(_f32)).astype(_f32)


def _mul3(a, s):
    return (a.astype(np.float64) * s).astype(_f32)


def _dot3(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]


def _len3(a):
    return np.sqrt(_dot3(a, a))


def _cross3(a, b):
    a = a.astype(_f32)
    b = b.astype(_f32)
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1], a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1).astype(_f32)


def _norm3(a):
    n = _len3(a)
    with np.errstate(divide="ignore"):
        inv = np.where(n > 0, 1.0 / np.where(n > 0, n, 1.0), 0.0)
    return np.where((n > 0)[..., None], _mul3(a, inv[..., None]), _f32(0))


def decode_placements(positions: np.ndarray, acts: np.ndarray, outputs: np.ndarray) -> np.ndarray:
    """src/neural/action_gpu.cu:1157-1188 without grid snapping: middle of the
    edge plus the clamped offset times the edge length, float positions."""
    p = positions.astype(_f32)
    a, b = p[acts[:, 0]], p[acts[:, 1]]
    middle = _mul3(_add3(a, b), 0.5)
    length = _len3(_sub3(a, b))
    y = np.minimum(_f32(1), np.maximum(_f32(-1), outputs[:, 1:4].astype(_f32)))
    return _add3(middle, _mul3(y, length[:, None]))


def pack_unorm16(value, low, extent) -> int:
    """src/neural/vertex_storage.hpp:24-29: (value - low) / extent * 65535 in
    double, rounded half up and clamped; zero extent encodes as zero."""
    if _f32(extent) == 0:
        return 0
    q = (float(_f32(value)) - float(_f32(low))) / float(_f32(extent)) * 65535.0
    return 0 if q <= 0 else 65535 if q >= 65535 else int(np.floor(q + 0.5))


def unpack_unorm16(v: int, low, extent) -> np.float32:
    """src/neural/vertex_storage.hpp:30-32: low + extent * (v / 65535), float."""
    return _f32(_f32(low) + _f32(extent) * (_f32(v) / _f32(65535)))


def grid_position(p: np.ndarray, low, extent) -> np.ndarray:
    """src/neural/action_gpu.cu:380-390: positions inside the box snap to its
    unorm16 grid, others are kept."""
    p = p.astype(_f32)
    low, extent = np.asarray(low, _f32), np.asarray(extent, _f32)
    if (p < low).any() or (p > (low + extent).astype(_f32)).any():
        return p
    return np.array([unpack_unorm16(pack_unorm16(p[k], low[k], extent[k]), low[k], extent[k]) for k in range(3)],
                    _f32)


def policy_placements(positions, normals, legal, mapping, y=None, supervise=False, quantization=None,
                      exact_bits=None) -> np.ndarray:
    """decode_placements / default_placements (src/neural/action_gpu.cu:
    1157-1200) for every legal action (from, to) with its wedge mapping
    (wedges, from0, to0, from1, to1): 9 floats, position then each
    destination wedge's unit normal."""
    p = positions.astype(_f32)
    out = np.zeros((len(legal), 9), _f32)

    def exact(v):
        return exact_bits is not None and bool((int(exact_bits[v // 32]) >> (v % 32)) & 1)

    for i, (u, v) in enumerate(legal):
        w = int(mapping[i, 0])
        pairs = [(int(mapping[i, 1]), int(mapping[i, 2])), (int(mapping[i, 3]), int(mapping[i, 4]))][:w]
        pos = p[v]
        if y is not None:
            middle = _mul3(_add3(p[u], p[v]), 0.5)
            length = _len3(_sub3(p[u], p[v]))
            off = np.minimum(_f32(1), np.maximum(_f32(-1), y[i, 3:6].astype(_f32)))
            pos = _add3(middle, _mul3(off, length))
            if quantization is not None and not any(exact(s) or exact(t) for s, t in pairs):
                pos = grid_position(pos, *quantization)
        out[i, :3] = pos
        if normals is not None:
            for j, (_, t) in enumerate(pairs):
                n = _norm3(normals[t].astype(_f32))
                if y is not None and supervise:
                    n = _norm3(_add3(n, y[i, 6 + 3 * j:9 + 3 * j].astype(_f32)))
                out[i, 3 + 3 * j:6 + 3 * j] = n
    return out


def teacher_placements(positions: np.ndarray, indices: np.ndarray, face_offsets, faces, acts: np.ndarray):
    """src/neural/action_gpu.cu:1304-1373 (positions, no grid snapping, no
    normals) for each action: 10 candidate positions (from, to, middle, the
    plane-quadric optimum of the faces around both endpoints, middle +/- a
    quarter edge along x, y, z) and their clamped offset targets."""
    p = positions.astype(_f32)
    idx = indices.reshape(-1, 3)
    out = np.zeros((len(acts), 10, 3), _f32)
    tgt = np.zeros((len(acts), 10, 3), _f32)
    for i, (u, v) in enumerate(acts):
        pf, pt = p[u], p[v]
        middle = _mul3(_add3(pf, pt), 0.5)
        scale = float(_len3(_sub3(pf, pt)))
        cand = [pf, pt, middle, middle]
        m = np.zeros((3, 4))
        for vert in (u, v):
            for k in range(face_offsets[vert], face_offsets[vert + 1]):
                a, b, c = p[idx[faces[k]]]
                cr = _cross3(_sub3(b, a), _sub3(c, a))
                n = _norm3(cr)
                length = float(_len3(cr))
                vv = n.astype(np.float64)
                d = float(_dot3(n, _sub3(a, middle)))
                for r in range(3):
                    for col in range(3):
                        m[r, col] += vv[r] * vv[col] * length
                    m[r, 3] += vv[r] * d * length
        solved = True
        for col in range(3):
            piv = col
            for r in range(col + 1, 3):
                if abs(m[r, col]) > abs(m[piv, col]):
                    piv = r
            if abs(m[piv, col]) < 1e-15:
                solved = False
                break
            m[[col, piv]] = m[[piv, col]]
            m[col, col:] = m[col, col:] / m[col, col]
            for r in range(3):
                if r != col:
                    m[r, col:] = m[r, col:] - m[r, col] * m[col, col:]
        if solved:
            cand[3] = _add3(middle, np.array([min(scale, max(-scale, m[r, 3])) for r in range(3)]).astype(_f32))
        for k in range(4, 10):
            delta = np.zeros(3, _f32)
            delta[(k - 4) // 2] = _f32((-0.25 if (k - 4) % 2 else 0.25) * scale)
            cand.append(_add3(middle, delta))
        with np.errstate(divide="ignore", invalid="ignore"):
            for k in range(10):
                out[i, k] = cand[k]
                off = _mul3(_sub3(cand[k], middle), 1.0 / scale if scale else np.inf)
                tgt[i, k] = np.fmin(_f32(1), np.fmax(_f32(-1), off))
    return out, tgt


def _cross32(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Vec3 cross product in float32, every product and difference rounded."""
    f = np.float32
    return np.stack([f(a[..., 1] * b[..., 2]) - f(a[..., 2] * b[..., 1]),
                     f(a[..., 2] * b[..., 0]) - f(a[..., 0] * b[..., 2]),
                     f(a[..., 0] * b[..., 1]) - f(a[..., 1] * b[..., 0])], -1).astype(np.float32)


def _dot64(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Vec3 dot product in double, summed x, y, z in order."""
    a, b = a.astype(np.float64), b.astype(np.float64)
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]


def scores(positions: np.ndarray, indices: np.ndarray, acts: np.ndarray, mode: int, outputs=None,
           diameter: float = 1.0) -> np.ndarray:
    """Ranking scores of src/neural/action_gpu.cu:809-844 as float32, in the
    reference's Vec3 arithmetic (float differences and cross products, double
    dot products and lengths; CPU ActionState::teacher_cost and the x[59]
    edge length of src/neural/action.cpp:355-395): 0 learned logit,
    1 constant, 3 -|dst - src| / diameter, 4 current plane: minus the sum over
    faces around src of (plane distance of dst / diameter)^2 times the face's
    |cross| / diameter^2, accumulated in face order."""
    n = len(acts)
    if mode == 0:
        return outputs[:, 0].astype(np.float32) + np.float32(0)
    if mode == 1:
        return np.zeros(n, np.float32)
    p = positions.astype(np.float32)
    scale = max(1e-20, diameter)
    src, dst = acts[:, 0].astype(np.int64), acts[:, 1].astype(np.int64)
    if mode == 3:
        d = (p[dst] - p[src]).astype(np.float32)
        e = (d.astype(np.float64) * (1 / scale)).astype(np.float32)
        return (-np.sqrt(_dot64(e, e)).astype(np.float32)) + np.float32(0)
    f = indices.astype(np.int64)
    _, _, face_offsets, faces, _ = topology(indices, len(positions))
    count = face_offsets[src + 1] - face_offsets[src]
    cost = np.zeros(n)
    for slot in range(int(count.max(initial=0))):
        a = np.nonzero(count > slot)[0]
        fi = faces[face_offsets[src[a]] + slot]
        P, Q, R = p[f[fi, 0]], p[f[fi, 1]], p[f[fi, 2]]
        normal = _cross32((Q - P).astype(np.float32), (R - P).astype(np.float32))
        weight = np.sqrt(_dot64(normal, normal))
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = np.where(weight > 0, 1 / np.where(weight > 0, weight, 1), 0.0)
        unit = (normal.astype(np.float64) * inv[:, None]).astype(np.float32)
        unit[~(weight > 0)] = 0
        distance = _dot64(unit, (p[dst[a]] - P).astype(np.float32)) / scale
        cost[a] += distance * distance * weight / (scale * scale)
    return (-cost).astype(np.float32) + np.float32(0)
