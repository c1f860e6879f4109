"""CPU oracle of one training step of the edge-graph scorer.

TEST INFRASTRUCTURE, NOT PRODUCT CODE: only tests/ import this module, as the
checker of the library's training entry point.

* `loss_fp64` restates the joint-scope loss of the reference's placement
  update (training/update_cuda.cu:256-330) for this network's four outputs
  [logit, placement x3]: pairwise hinge between preferred and other valid
  entries of each pool (margin m, each pair weighted 1 / pairs over the
  batch, update_cuda.cu:270-291), an L2 penalty of weight penalty / valid / 3
  on the logit of valid entries (update_cuda.cu:293-299, the h = 0 term), and
  smooth-L1 (delta 1) on the three placement channels of placed entries,
  normalised by 3 x placed (update_cuda.cu:311-322).
* `backward_fp64` is the float64 gradient of that loss through the network
  (mean aggregation transposed over the CSR graph as given, without assuming
  it is symmetric).
* `adamw_fp32` restates the clip and AdamW update (update_cuda.cu:358-393)
  operation for operation in float32.
"""
from __future__ import annotations

import numpy as np

from scorer_oracle import LAYERS, SHAPES, split_weights

VALID, PREFERRED, PLACED = 1, 2, 4


def pools_counts(labels: np.ndarray, pool: int) -> tuple[int, int, int]:
    """(pairs, valid, placed) over the batch (update_cuda.cu:216-254)."""
    pairs = valid = 0
    for p0 in range(0, len(labels), pool):
        lab = labels[p0:p0 + pool]
        v = int(np.count_nonzero(lab & VALID))
        p = int(np.count_nonzero((lab & (VALID | PREFERRED)) == (VALID | PREFERRED)))
        pairs += p * (v - p)
        valid += v
    placed = int(np.count_nonzero(labels & PLACED))
    return pairs, valid, placed


def loss_fp64(y: np.ndarray, labels: np.ndarray, targets, pool: int, margin=1.0, penalty=1e-4,
              placement_weight=1.0):
    """Loss and dL/dy (float64) of a batch of outputs y (count x 4)."""
    y = y.astype(np.float64)
    n = len(labels)
    pairs, valid, placed = pools_counts(labels, pool)
    dy = np.zeros((n, 4))
    loss = 0.0
    inv_pairs = 1.0 / max(1, pairs)
    w_pen = penalty / max(1, valid) / 3
    for p0 in range(0, n, pool):
        idx = np.arange(p0, min(n, p0 + pool))
        ok = (labels[idx] & VALID) != 0
        pref = ok & ((labels[idx] & PREFERRED) != 0)
        other = ok & ~pref
        for i in idx[pref]:
            for k in idx[other]:
                v = margin + y[k, 0] - y[i, 0]
                if v > 0:
                    loss += v * inv_pairs
                    dy[i, 0] -= inv_pairs
                    dy[k, 0] += inv_pairs
        for i in idx[ok]:
            loss += w_pen * y[i, 0] ** 2
            dy[i, 0] += 2 * w_pen * y[i, 0]
    if placed:
        inv = placement_weight / (placed * 3.0)
        for i in np.flatnonzero(labels & PLACED):
            d = y[i, 1:4] - np.asarray(targets, np.float64)[i]
            a = np.abs(d)
            loss += float(np.sum(np.where(a < 1, 0.5 * d * d, a - 0.5))) * inv
            dy[i, 1:4] = np.clip(d, -1, 1) * inv
    return loss, dy


def forward_cache(x, offsets, neighbors, batch, weights):
    layers = [(W.astype(np.float64), b.astype(np.float64)) for W, b in split_weights(weights)]
    deg = np.maximum(np.diff(offsets), 1).astype(np.float64)
    src = np.repeat(np.arange(len(offsets) - 1), np.diff(offsets))
    hs, zs = [x.astype(np.float64)], []
    for W, b in layers[:LAYERS]:
        h = hs[-1]
        agg = np.zeros_like(h)
        np.add.at(agg, src, h[neighbors])
        z = np.concatenate([h, agg / deg[:, None]], 1)
        zs.append(z)
        hs.append(np.maximum(z @ W.T + b, 0.0))
    ze = np.concatenate([hs[-1][batch[:, 0]], hs[-1][batch[:, 1]]], 1)
    We, be = layers[3]
    Wo, bo = layers[4]
    g = np.maximum(ze @ We.T + be, 0.0)
    y = g @ Wo.T + bo
    return dict(layers=layers, hs=hs, zs=zs, ze=ze, g=g, y=y, deg=deg, src=src)


def loss_of(weights, x, offsets, neighbors, batch, labels, targets, pool, **kw) -> float:
    c = forward_cache(x, offsets, neighbors, batch, weights.astype(np.float64))
    return loss_fp64(c["y"], labels, targets, pool, **kw)[0]


def backward_fp64(x, offsets, neighbors, batch, weights, labels, targets, pool, plant: str | None = None, **kw):
    """(loss, flat float64 gradient in the weight-buffer layout).

    `plant` swaps in a deliberate mistake for the tests that show the
    finite-difference check catches it: "degree" divides the aggregation
    transpose by the receiving vertex's degree, "mask" drops the edge
    layer's ReLU mask.
    """
    c = forward_cache(x, offsets, neighbors, batch, weights)
    loss, dy = loss_fp64(c["y"], labels, targets, pool, **kw)
    layers, hs, zs = c["layers"], c["hs"], c["zs"]
    grads = [None] * len(layers)
    We, _ = layers[3]
    Wo, _ = layers[4]
    grads[4] = (dy.T @ c["g"], dy.sum(0))
    dg = dy @ Wo
    dpe = dg if plant == "mask" else dg * (c["g"] > 0)
    grads[3] = (dpe.T @ c["ze"], dpe.sum(0))
    dze = dpe @ We
    d = hs[-1].shape[1]
    dh = np.zeros_like(hs[-1])
    np.add.at(dh, batch[:, 0], dze[:, :d])
    np.add.at(dh, batch[:, 1], dze[:, d:])
    deg, src = c["deg"], c["src"]
    for l in range(LAYERS - 1, -1, -1):
        W, _ = layers[l]
        dpre = dh * (hs[l + 1] > 0)
        grads[l] = (dpre.T @ zs[l], dpre.sum(0))
        if l == 0:
            break
        dz = dpre @ W
        w = hs[l].shape[1]
        dh = dz[:, :w].copy()
        # agg[v] = sum over u in N(v) of h[u] / deg(v): h[u] receives dz_agg[v] / deg(v).
        scale = deg[neighbors] if plant == "degree" else deg[src]
        np.add.at(dh, neighbors, dz[src, w:] / scale[:, None])
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def layer_slices() -> list[tuple[slice, slice]]:
    """(matrix, bias) slices of every layer in the flat weight buffer."""
    out, at = [], 0
    for o, i in SHAPES:
        out.append((slice(at, at + o * i), slice(at + o * i, at + o * i + o)))
        at += o * (i + 1)
    return out


def clip_and_corrections(grad_norm: np.float32, step: int, lr=1e-3, max_norm=1.0):
    """update_cuda.cu:366-375: clip factor and bias corrections."""
    f = np.float32
    clip = np.minimum(f(1), f(max_norm) / (f(grad_norm) + f(1e-6)))
    c1 = f(np.float64(f(lr)) / (1 - 0.9 ** (float(step) + 1)))
    c2 = f(np.sqrt(1 - 0.999 ** (float(step) + 1)))
    return clip, c1, c2


def adamw_fp32(value, grad, mean, variance, clip, c1, c2, lr=1e-3, decay=1e-4):
    """update_cuda.cu:378-393 in float32: returns (value, mean, variance)."""
    f = np.float32
    value, grad, mean, variance = (np.asarray(a, f) for a in (value, grad, mean, variance))
    g = grad * f(clip)
    m = mean * f(0.9) + g * f(0.1)
    v = variance * f(0.999) + (g * g) * f(0.001)
    shrink = f(1) - f(lr) * f(decay)
    out = value * shrink - f(c1) * m / (np.sqrt(v) / f(c2) + f(1e-8))
    return out.astype(f), m.astype(f), v.astype(f)
