"""Scorer training step: oracle pins (CPU) and HIP parity (GPU).

The FP64 backward of oracle/train_oracle.py is pinned by central finite
differences of its own loss, with planted mistakes (a wrong aggregation
degree, a dropped ReLU mask) shown to fail that check; the loss by a
hand-evaluated pool; AdamW by the closed form of its first step. GPU checks:
gradients of the FP32 and BF16 steps against FP64 (norm-relative, per layer),
the AdamW update bit-exact against the float32 restatement over two steps,
the loss falling over repeated steps, and rejected inputs.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "oracle"))

import scorer_oracle as so  # noqa: E402
import train_oracle as to  # noqa: E402
from blitzremesher_amd import meshes  # noqa: E402
from blitzremesher_amd.audit import bounds  # noqa: E402

VALID, PREFERRED, PLACED = to.VALID, to.PREFERRED, to.PLACED


def random_batch(actions: int, count: int, seed: int, placed=0.5):
    rng = np.random.default_rng(seed)
    ids = rng.choice(actions, count, replace=False).astype(np.uint32)
    valid = rng.random(count) < 0.85
    pref = valid & (rng.random(count) < 0.25)
    plc = rng.random(count) < placed
    labels = (valid * VALID | pref * PREFERRED | plc * PLACED).astype(np.uint8)
    targets = rng.uniform(-1.6, 1.6, (count, 3)).astype(np.float32)
    return ids, labels, targets


def small_case(seed=0):
    mesh = meshes.icosphere(1, 0.05, 3)
    c, r = bounds(mesh)
    offsets, neighbors, *_ = so.topology(mesh.indices, len(mesh.positions))
    acts = so.actions(offsets, neighbors)
    x = so.features(mesh.positions, mesh.indices, c, 2 * r).astype(np.float32)
    ids, labels, targets = random_batch(len(acts), 60, seed)
    return x, offsets, neighbors, acts[ids], labels, targets


def weights_with_bias(seed=4):
    from blitzremesher_amd.scorer import he_weights
    w = he_weights(seed).astype(np.float64)
    rng = np.random.default_rng(seed)
    for _, bias in to.layer_slices():
        w[bias] = rng.normal(0, 0.05, bias.stop - bias.start)
    return w


def finite_difference_errors(plant=None, coords=80, eps=1e-6):
    x, offsets, neighbors, batch, labels, targets = small_case()
    w = weights_with_bias()
    _, grad = to.backward_fp64(x, offsets, neighbors, batch, w, labels, targets, 15, plant=plant)
    rng = np.random.default_rng(7)
    picks = []
    for mat, bias in to.layer_slices():
        picks += list(rng.integers(mat.start, mat.stop, coords // 5))
        picks.append(bias.start)
    errs = []
    for i in picks:
        wp, wm = w.copy(), w.copy()
        wp[i] += eps
        wm[i] -= eps
        fd = (to.loss_of(wp, x, offsets, neighbors, batch, labels, targets, 15)
              - to.loss_of(wm, x, offsets, neighbors, batch, labels, targets, 15)) / (2 * eps)
        errs.append(abs(fd - grad[i]) / (1e-4 + abs(fd)))
    return np.array(errs)


def test_backward_oracle_matches_finite_differences():
    errs = finite_difference_errors()
    assert errs.max() < 1e-4, errs.max()


@pytest.mark.parametrize("plant", ["degree", "mask"])
def test_finite_differences_catch_planted_mistakes(plant):
    assert finite_difference_errors(plant).max() > 1e-2


def test_loss_closed_form():
    # One pool of four: entry 0 preferred (0.2), entries 1, 2 other valid
    # (0.5, -2.0), entry 3 invalid; entry 1 placed with delta (0.5, -2, 0).
    y = np.array([[0.2, 0, 0, 0], [0.5, 0.5, -1.0, 0.25], [-2.0, 0, 0, 0], [9.0, 0, 0, 0]])
    labels = np.array([VALID | PREFERRED, VALID | PLACED, VALID, 0], np.uint8)
    targets = np.array([[0, 0, 0], [0, 1.0, 0.25], [0, 0, 0], [0, 0, 0]])
    loss, dy = to.loss_fp64(y, labels, targets, 4, margin=1.0, penalty=1e-4)
    hinge = (1 + 0.5 - 0.2) / 2  # pairs = 1 x 2; the pair with -2.0 is inactive
    penalty = 1e-4 / 3 / 3 * (0.2 ** 2 + 0.5 ** 2 + 2.0 ** 2)
    placement = (0.5 * 0.5 ** 2 + (2 - 0.5) + 0) / 3
    assert loss == pytest.approx(hinge + penalty + placement, rel=1e-12)
    w = 1e-4 / 9
    assert dy[0, 0] == pytest.approx(-0.5 + 2 * w * 0.2, rel=1e-12)
    assert dy[1, 0] == pytest.approx(0.5 + 2 * w * 0.5, rel=1e-12)
    assert dy[2, 0] == pytest.approx(2 * w * -2.0, rel=1e-12)
    assert np.allclose(dy[1, 1:], [0.5 / 3, -1 / 3, 0])
    assert not dy[3].any()


def test_adamw_first_step_closed_form():
    # From zero moments the bias corrections cancel the moment decay:
    # w1 = w0 (1 - lr decay) - lr g / (|g| + 1e-8).
    rng = np.random.default_rng(1)
    w0 = rng.normal(0, 1, 1000).astype(np.float32)
    g = rng.normal(0, 1e-2, 1000).astype(np.float32)
    clip, c1, c2 = to.clip_and_corrections(np.float32(0.5), 0, lr=1e-3, max_norm=1.0)
    assert clip == 1
    w1, m, v = to.adamw_fp32(w0, g, np.zeros(1000), np.zeros(1000), clip, c1, c2, lr=1e-3, decay=1e-4)
    g64 = g.astype(np.float64)
    want = w0.astype(np.float64) * (1 - 1e-7) - 1e-3 * g64 / (np.abs(g64) + 1e-8)
    assert np.abs(w1 - want).max() < 1e-6
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 1e-3 * g.astype(np.float64) ** 2, rtol=1e-5)
    # Clipping scales the gradient by max_norm / norm.
    clip, *_ = to.clip_and_corrections(np.float32(4.0), 0, max_norm=1.0)
    assert clip == pytest.approx(0.25, rel=1e-6)


def test_pool_counts():
    labels = np.array([3, 1, 1, 0, 3, 3, 1, 5], np.uint8)
    assert to.pools_counts(labels, 4) == (1 * 2 + 2 * 2, 3 + 4, 1)


# ----------------------------------------------------------------------- GPU

@pytest.fixture(scope="module")
def scorer():
    from blitzremesher_amd.scorer import Scorer, he_weights
    s = Scorer(he_weights(4))
    yield s
    s.close()


def gpu_case(scorer, mesh, count, seed):
    c, r = bounds(mesh)
    g = scorer.graph(mesh, c, r)
    offsets, neighbors, *_ = so.topology(mesh.indices, len(mesh.positions))
    acts = so.actions(offsets, neighbors)
    ids, labels, targets = random_batch(len(acts), count, seed)
    return g, offsets, neighbors, acts, ids, labels, targets


def layer_errors(got, want):
    out = []
    for mat, bias in to.layer_slices():
        for sl in (mat, bias):
            ref = want[sl]
            out.append(np.linalg.norm(got[sl] - ref) / max(np.linalg.norm(ref), 1e-12))
    return np.array(out)


# FP32: activations, masks and sums in float32 against float64; measured
# norm-relative gradient errors per layer are ~7e-7. BF16: activations and
# weights rounded to 8-bit mantissas in the forward pass that the FP32
# backward differentiates; measured 2e-2..5e-2 per layer.
FP32_GRAD_TOLERANCE = 1e-5
BF16_GRAD_TOLERANCE = 0.1

TRAIN_CASES = {
    "icosphere": lambda: meshes.icosphere(3, 0.02, 1),
    "torus": lambda: meshes.torus(1.0, 0.3, 40, 16, 0.01, 3),
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(TRAIN_CASES))
@pytest.mark.parametrize("precision", [0, 1])
def test_gpu_gradients_match_fp64(scorer, case, precision):
    check_gradients(scorer, case, precision, 512)


@pytest.mark.gpu
@pytest.mark.parametrize("count", [1, 17, 509])
def test_gpu_gradients_match_fp64_ragged_batches(scorer, count):
    # Batch rows that fill neither a 16-row GEMM step nor a pool of 32.
    check_gradients(scorer, "torus", 0, count)


def check_gradients(scorer, case, precision, count):
    from blitzremesher_amd.scorer import TrainSettings, he_weights
    w = he_weights(4)
    scorer.set_weights(w)
    g, offsets, neighbors, acts, ids, labels, targets = gpu_case(scorer, TRAIN_CASES[case](), count, 11)
    s = TrainSettings(pool=32, precision=precision)
    r = scorer.train_step(g, ids, labels, targets, s)
    grad = scorer.gradients()
    loss, want = to.backward_fp64(g.features(), offsets, neighbors, acts[ids], w.astype(np.float64), labels,
                                  targets, 32)
    errs = layer_errors(grad.astype(np.float64), want)
    print(f"{case} precision {precision}: loss {r.loss:.6f} vs {loss:.6f}; grad norm {r.grad_norm:.4e} vs "
          f"{np.linalg.norm(want):.4e}; layer errors {np.array2string(errs, precision=2)}")
    assert r.applied and r.step == 1
    tol = FP32_GRAD_TOLERANCE if precision == 0 else BF16_GRAD_TOLERANCE
    assert abs(r.loss - loss) <= tol * max(1.0, abs(loss))
    assert errs.max() <= tol
    assert abs(r.grad_norm - np.linalg.norm(grad.astype(np.float64))) <= 1e-5 * r.grad_norm


@pytest.mark.gpu
def test_gpu_adamw_matches_float32_restatement(scorer):
    from blitzremesher_amd.scorer import TrainSettings, he_weights
    w0 = he_weights(5)
    scorer.set_weights(w0)
    g, offsets, neighbors, acts, ids, labels, targets = gpu_case(scorer, TRAIN_CASES["torus"](), 256, 3)
    s = TrainSettings(pool=16, precision=0, lr=2e-3, decay=1e-3, max_norm=0.05)
    mean = np.zeros_like(w0)
    var = np.zeros_like(w0)
    w = w0
    for step in range(2):
        r = scorer.train_step(g, ids, labels, targets, s)
        grad = scorer.gradients()
        clip, c1, c2 = to.clip_and_corrections(np.float32(r.grad_norm), step, lr=s.lr, max_norm=s.max_norm)
        w, mean, var = to.adamw_fp32(w, grad, mean, var, clip, c1, c2, lr=s.lr, decay=s.decay)
        got = scorer.weights()
        assert r.step == step + 1
        assert np.array_equal(got, w), np.abs(got - w).max()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_gpu_training_reduces_loss(scorer, precision):
    from blitzremesher_amd.scorer import TrainSettings, he_weights
    scorer.set_weights(he_weights(6))
    g, offsets, neighbors, acts, ids, labels, targets = gpu_case(scorer, TRAIN_CASES["icosphere"](), 1024, 5)
    s = TrainSettings(pool=32, precision=precision, lr=2e-3)
    losses = [scorer.train_step(g, ids, labels, targets, s).loss for _ in range(40)]
    print(f"precision {precision}: loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.7 * losses[0]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_gpu_training_is_deterministic(scorer, precision):
    # Every reduction (loss, weight-gradient slices, edge-gradient gathers,
    # gradient norm) adds its partials in a fixed order: two runs from the
    # same weights and optimizer state end bit-identical. The torus has
    # 12,000 vertices, so the vertex-layer weight gradients are summed over
    # many row slices.
    from blitzremesher_amd.scorer import TrainSettings, he_weights
    g, offsets, neighbors, acts, ids, labels, targets = gpu_case(
        scorer, meshes.torus(1.0, 0.3, 200, 60, 0.01, 9), 4096, 21)
    s = TrainSettings(pool=32, precision=precision, lr=2e-3)
    runs = []
    for _ in range(2):
        scorer.set_weights(he_weights(8))
        scorer.reset_optimizer()
        results = [scorer.train_step(g, ids, labels, targets, s) for _ in range(4)]
        runs.append(([(r.loss, r.grad_norm, r.step) for r in results], scorer.gradients(), scorer.weights()))
    (res_a, grad_a, w_a), (res_b, grad_b, w_b) = runs
    assert res_a == res_b
    assert np.array_equal(grad_a.view(np.uint32), grad_b.view(np.uint32))
    assert np.array_equal(w_a.view(np.uint32), w_b.view(np.uint32))


@pytest.mark.gpu
def test_gpu_train_rejects_bad_batches(scorer):
    from blitzremesher_amd._lib import BlitzError
    from blitzremesher_amd.scorer import TrainSettings
    g, offsets, neighbors, acts, ids, labels, targets = gpu_case(scorer, TRAIN_CASES["icosphere"](), 64, 1)
    bad = ids.copy()
    bad[3] = len(acts)
    with pytest.raises(BlitzError):
        scorer.train_step(g, bad, labels, targets)
    with pytest.raises(BlitzError):
        scorer.train_step(g, ids, labels, targets, TrainSettings(pool=0))
    with pytest.raises(BlitzError):
        scorer.train_step(g, ids, labels, targets, TrainSettings(pool=300))
    with pytest.raises(BlitzError):
        scorer.train_step(g, ids, labels | PLACED, None)
    with pytest.raises(BlitzError):
        scorer.train_step(g, ids, labels, targets, TrainSettings(lr=float("nan")))
