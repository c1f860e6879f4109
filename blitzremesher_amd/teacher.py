"""Training labels for the edge-graph scorer, made by the classical order and
the chain's four audits, and batches that span many meshes.

A pool is the reference's teacher query of one state
(training/runtime_teacher.hpp:125-199): up to 16 legal collapses, each
audited under the four gates of the level (source and adjacent, each under
the search and the audit settings), labelled with the PlacementLabel bits
(src/neural/placement.hpp:42-51), and the collapses that pass all four with
the fewest faces and then the smallest margin marked preferred
(preferred_actions, training/teacher_labels.hpp:16-36). Rows are chosen as
teacher_selection chooses them (src/neural/action_gpu.cu:1216-1235): the
first half from the classical current-plane order, the rest distinct random
legal actions.

All pools of a mesh are queried at its source state, where the previous
level of detail is the source itself: pool p takes its classical half from
ranks 8p.. of the order, the head a classical rollout reaches after p steps
of eight collapses, so the pools of a mesh walk down its order. Each row is
audited as a patch of the bound source (every collapse onto an existing
vertex). The placement target of a row that passes all four gates is the
teacher solve of that collapse: the plane-quadric optimum of the faces
around both endpoints as an offset from the edge middle in edge lengths
(teacher_placements, src/neural/action_gpu.cu:1304-1373); the reference
audits its teacher placements one by one, this target is not audited.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .audit import EvalSettings, Measurement
from .scorer import PLACED, PREFERRED, VALID

# PlacementLabel (src/neural/placement.hpp:42-51).
SOURCE_PASS, ADJACENT_PASS, PLACEMENT_PREFERRED, SOURCE_KNOWN, ADJACENT_KNOWN = 1, 2, 4, 8, 16
POOL = 16  # teacher pools hold at most 16 rows (training/teacher_labels.hpp:18)


@dataclass
class Choice:
    """TeacherChoice (training/teacher_labels.hpp:9-13)."""
    faces: int = 0
    margin: float = 0.0
    confirmed: bool = False


def preferred_actions(choices: list[Choice], minimum_faces: int = 0) -> int:
    """Bit mask of the preferred choices (training/teacher_labels.hpp:16-36):
    among confirmed choices with a finite margin and at least
    `minimum_faces` faces, those with the fewest faces and, among them, the
    smallest margin; exact ties all preferred."""
    if len(choices) > 16:
        raise ValueError("teacher choice mask capacity")
    faces, margin, result = None, math.inf, 0
    for i, c in enumerate(choices):
        if not c.confirmed or not math.isfinite(c.margin) or c.faces < minimum_faces:
            continue
        if faces is None or c.faces < faces or (c.faces == faces and c.margin < margin):
            faces, margin, result = c.faces, c.margin, 0
        if c.faces == faces and c.margin == margin:
            result |= 1 << i
    return result


def nonfinite(m: Measurement) -> bool:
    """audit_measurement_nonfinite (src/neural/audit_measurement.hpp:8-13)."""
    return (math.isnan(m.error) or math.isnan(m.coverage) or math.isnan(m.coverage_upper)
            or not math.isfinite(m.changed_area) or not math.isfinite(m.normal_degrees)
            or (m.passed and not (math.isfinite(m.error) and math.isfinite(m.coverage)
                                  and math.isfinite(m.coverage_upper))))


def audit_known(m: Measurement, e: EvalSettings) -> bool:
    """action_audit_known (src/neural/action.cpp:21-26): a complete pass, or
    a failure witnessed by an error or changed area over its limit."""
    if m.cancelled or m.resource_limited or nonfinite(m):
        return False
    return (m.complete and m.passed) or (m.views_evaluated > 0 and not m.passed
                                         and (m.error > e.limit or m.changed_area > e.max_changed_area))


def _ratio(value: float, limit: float) -> float:
    return value / limit if limit > 0 else (0.0 if value == 0 else math.inf)


def row_label(x: Measurement, y: Measurement, sx: Measurement, sy: Measurement,
              gates: tuple[EvalSettings, EvalSettings, EvalSettings, EvalSettings], faces: int) -> tuple[int, Choice]:
    """Label bits and teacher choice of one audited row
    (training/runtime_teacher.hpp:152-184). gates = (source search, adjacent
    search, source audit, adjacent audit); x, y are the source and adjacent
    audits, sx, sy the searches."""
    search_a, search_b, a, b = gates
    known_x = audit_known(x, a) and audit_known(sx, search_a)
    known_y = audit_known(y, b) and audit_known(sy, search_b)
    label = ((SOURCE_KNOWN if known_x else 0) | (ADJACENT_KNOWN if known_y else 0)
             | (SOURCE_PASS if known_x and x.passed and sx.passed else 0)
             | (ADJACENT_PASS if known_y and y.passed and sy.passed else 0))
    choice = Choice()
    if known_x and known_y and x.passed and y.passed and sx.passed and sy.passed:
        margin = max(x.error / a.limit, y.error / b.limit, sx.error / search_a.limit, sy.error / search_b.limit,
                     _ratio(x.changed_area, a.max_changed_area), _ratio(y.changed_area, b.max_changed_area),
                     _ratio(sx.changed_area, search_a.max_changed_area),
                     _ratio(sy.changed_area, search_b.max_changed_area))
        choice = Choice(faces, margin, x.complete and y.complete and sx.complete and sy.complete)
    return label, choice


def level_gates(settings, center, radius, level: int = 1):
    """The four EvalSettings of a chain level in gate order (source search,
    adjacent search, source audit, adjacent audit), as eval_config builds them
    (src/chain.cpp:179-199) from the level's schedule entry."""
    from .chain import schedule
    pixels, transition, source = schedule(settings, center, radius)[level]
    t = 0.0 if settings.levels == 2 else (level - 1) / (settings.levels - 2)

    def at(curve):  # piecewise linear in the given point order, clamped at the ends
        pts = list(curve)
        if t <= pts[0][0]:
            return pts[0][1]
        for (x0, y0), (x1, y1) in zip(pts, pts[1:]):
            if t <= x1:
                u = (t - x0) / (x1 - x0)
                return y0 * (1 - u) + y1 * u
        return pts[-1][1]

    def one(audit: bool, limit: float) -> EvalSettings:
        return EvalSettings(profile=settings.profile, screen_size=pixels, limit=limit,
                            max_changed_area=settings.max_changed_area if audit else 1.0,
                            normal_weight=settings.normal_weight * at(settings.normal_importance),
                            color_weight=settings.color_weight * at(settings.attribute_importance),
                            material_weight=settings.material_weight * at(settings.attribute_importance),
                            orthographic=settings.audit_orthographic if audit else settings.search_orthographic,
                            perspective=settings.audit_perspective if audit else settings.search_perspective,
                            rotation_seed=settings.audit_seed if audit else settings.search_seed,
                            supersample=settings.audit_supersample if audit else settings.search_supersample,
                            max_supersample=settings.max_supersample)
    return one(False, source), one(False, transition), one(True, source), one(True, transition)


def pool_rows(order: np.ndarray, pools: int, seed: int) -> np.ndarray:
    """(pools, 16) legal-action indices: row i < 8 of pool p is order[8p + i],
    rows 8.. distinct uniform random legal actions not already in the pool
    (teacher_selection, src/neural/action_gpu.cu:1216-1235, with this
    package's own seeded generator)."""
    n = len(order)
    half = (POOL + 1) // 2
    if n < POOL or pools < 1 or pools * half > n:
        raise ValueError("not enough legal actions for the pools")
    rng = np.random.default_rng(seed)
    out = np.empty((pools, POOL), np.uint32)
    for p in range(pools):
        head = order[p * half:(p + 1) * half]
        taken = set(int(v) for v in head)
        rest = []
        while len(rest) < POOL - half:
            v = int(rng.integers(n))
            if v not in taken:
                taken.add(v)
                rest.append(v)
        out[p, :half] = head
        out[p, half:] = rest
    return out


@dataclass
class MeshLabels:
    """The labelled pools of one mesh: rows in pool order (16 per pool)."""
    legal: np.ndarray        # (rows,) legal-action indices of the mesh's ActionSet
    pairs: np.ndarray        # (rows, 2) collapse (from, to)
    label: np.ndarray        # (rows,) PlacementLabel bits
    faces: np.ndarray        # (rows,) candidate faces
    margin: np.ndarray       # (rows,) teacher margin (inf when not confirmed)
    included: np.ndarray     # (rows,) rows before the pool's first unknown, inclusive
    fallbacks: int = 0


def label_mesh(audit, mesh, center, radius, gates, pools: int, seed: int, target_triangles: int = 0,
               uv=None) -> MeshLabels:
    """Audits up to `pools` teacher pools of `mesh` (at most one per 16
    legal actions) at its source state under the four gates and labels them
    (training/runtime_teacher.hpp:125-199)."""
    from .actions import ActionSet
    search_a, search_b, a, b = gates
    with ActionSet(mesh, uv) as acts:
        order = acts.rank(4)
        pools = min(pools, len(order) // POOL)
        if pools < 1:
            raise ValueError("mesh has fewer than 16 legal actions")
        rows = pool_rows(order, pools, seed).reshape(-1)
        patches = acts.patches(rows)
        pairs = acts.legal[rows].copy()
    search = audit.bind(mesh, center, radius, search_a)
    full = audit.bind(mesh, center, radius, a)
    fallbacks = 0
    ms = []
    for binding, gate in ((search, search_a), (search, search_b), (full, a), (full, b)):
        m, f = audit.evaluate_patches(binding, patches, gate)
        ms.append(m)
        fallbacks += f
    search.release()
    full.release()
    sx, sy, x, y = ms
    removed = np.diff(patches.removed_offsets).astype(np.int64)
    added = np.diff(patches.added_offsets).astype(np.int64)
    faces = mesh.triangles - removed + added
    n = len(rows)
    label = np.zeros(n, np.uint8)
    margin = np.full(n, np.inf)
    included = np.zeros(n, bool)
    for p in range(pools):
        choices = []
        for i in range(p * POOL, (p + 1) * POOL):
            bits, choice = row_label(x[i], y[i], sx[i], sy[i], gates, int(faces[i]))
            label[i] = bits
            included[i] = True
            if choice.confirmed:
                margin[i] = choice.margin
            choices.append(choice)
            if not (bits & SOURCE_KNOWN and bits & ADJACENT_KNOWN):
                break  # an unknown observation ends the pool
        mask = preferred_actions(choices, target_triangles)
        for k in range(len(choices)):
            if mask >> k & 1:
                label[p * POOL + k] |= PLACEMENT_PREFERRED
    return MeshLabels(rows, pairs, label, faces, margin, included, fallbacks)


def training_labels(labels: MeshLabels) -> tuple[np.ndarray, np.ndarray]:
    """Scorer training bits of each row (loss_gradient,
    training/update_cuda.cu:159-205: a row takes part when its source verdict
    is known, and ranks as preferred when PlacementPreferred is set) plus the
    placement bit for confirmed rows; and the confirmed mask."""
    lab = labels.label
    use = labels.included & ((lab & SOURCE_KNOWN) != 0)
    confirmed = use & np.isfinite(labels.margin)
    bits = (np.where(use, VALID, 0) | np.where(use & ((lab & PLACEMENT_PREFERRED) != 0), PREFERRED, 0)
            | np.where(confirmed, PLACED, 0)).astype(np.uint8)
    return bits, confirmed


__all__ = ["Choice", "MeshLabels", "POOL", "audit_known", "label_mesh", "level_gates", "pool_rows",
           "preferred_actions", "row_label", "training_labels"]


@dataclass
class Dataset:
    """Labelled pools of many meshes over one joined scorer graph: rows in
    pools of 16, action ids of the joined graph."""
    graph: object            # scorer Graph joining every mesh's graph
    actions: np.ndarray      # (rows,) action ids of the joined graph
    bits: np.ndarray         # (rows,) VALID / PREFERRED / PLACED
    targets: np.ndarray      # (rows, 3) placement targets (zero where not placed)
    pool_mesh: np.ndarray    # (pools,) mesh of each pool
    meshes: list             # MeshLabels per mesh
    fallbacks: int = 0

    @property
    def pools(self) -> int:
        return len(self.pool_mesh)


def graph_action_ids(graph, pairs: np.ndarray) -> np.ndarray:
    """Action ids of the directed collapses `pairs` (from, to) in `graph`."""
    acts = graph.actions_array()
    keys = acts[:, 0].astype(np.uint64) << np.uint64(32) | acts[:, 1].astype(np.uint64)
    want = pairs[:, 0].astype(np.uint64) << np.uint64(32) | pairs[:, 1].astype(np.uint64)
    idx = np.searchsorted(keys, want)
    if not np.array_equal(keys[np.minimum(idx, len(keys) - 1)], want):
        raise ValueError("collapse missing from the scorer graph")
    return idx.astype(np.uint32)


def build_dataset(audit, scorer, meshes, settings, pools_per_mesh: int, seed: int = 5) -> Dataset:
    """Labels `pools_per_mesh` teacher pools of every mesh (fewer on meshes
    with too few legal actions) at its source state under the first level's
    gates of `settings`, and joins the meshes' scorer graphs so that batches
    span meshes. Targets: the teacher solve of each confirmed row."""
    from .audit import bounds
    parts, labels, ids, bits, targets, pool_mesh = [], [], [], [], [], []
    fallbacks = 0
    for k, mesh in enumerate(meshes):
        center, radius = bounds(mesh)
        gates = level_gates(settings, center, radius)
        lab = label_mesh(audit, mesh, center, radius, gates, pools_per_mesh, seed * 7919 + k)
        pools = len(lab.legal) // POOL
        fallbacks += lab.fallbacks
        g = scorer.graph(mesh, center, radius)
        row_ids = graph_action_ids(g, lab.pairs)
        b, confirmed = training_labels(lab)
        t = np.zeros((len(row_ids), 3), np.float32)
        if confirmed.any():
            _, tgt = scorer.teacher(g, row_ids[confirmed])
            t[confirmed] = tgt[:, 3]
        parts.append(g)
        labels.append(lab)
        ids.append(row_ids)
        bits.append(b)
        targets.append(t)
        pool_mesh.append(np.full(pools, k, np.int32))
    joined = scorer.concat(parts)
    base = 0
    for k, g in enumerate(parts):
        ids[k] = ids[k] + np.uint32(base)
        base += g.actions
        g.release()
    return Dataset(joined, np.concatenate(ids), np.concatenate(bits), np.concatenate(targets),
                   np.concatenate(pool_mesh), labels, fallbacks)


def sample_batch(data: Dataset, pools: np.ndarray, batch: int, rng: np.random.Generator):
    """A batch of `batch` rows: batch / 16 pools drawn without replacement
    from `pools` (pool indices of the dataset), whole pools in order."""
    take = rng.choice(pools, batch // POOL, replace=False)
    rows = (take[:, None] * POOL + np.arange(POOL)).reshape(-1)
    return data.actions[rows], data.bits[rows], data.targets[rows]


def top1_preferred(data: Dataset, logits: np.ndarray, pools: np.ndarray) -> float:
    """Fraction of the given pools, among those with a preferred and a
    non-preferred valid row, whose highest-logit valid row is preferred."""
    hit = total = 0
    for p in pools:
        rows = slice(p * POOL, (p + 1) * POOL)
        b = data.bits[rows]
        valid = (b & VALID) != 0
        pref = valid & ((b & PREFERRED) != 0)
        if not pref.any() or pref.sum() == valid.sum():
            continue
        s = np.where(valid, logits[data.actions[rows]], -np.inf)
        total += 1
        hit += bool(pref[int(np.argmax(s))])
    return hit / max(total, 1)
