"""LOD-chain cases shared by the golden generator and the chain tests.

Each case is a seeded mesh and the settings of one chain: coverage profile,
explicit reuse output, proposals from the source, the collapse executor
ranked by current plane (4) or shortest edge (3). Sizes are chosen so that the
reference's own CPU chain (oracle/_ref) finishes in one to two minutes.
"""
from __future__ import annotations

from blitzremesher_amd import meshes
from blitzremesher_amd.chain import ChainOptions, ChainSettings

SEARCH_SEED, AUDIT_SEED = 0xB1172024, 0xB1172026


def chain_cases() -> dict:
    """name -> (mesh, ChainSettings, ChainOptions)."""
    cases = {}
    cases["ico2_plane"] = (
        meshes.icosphere(2, 0.02, 1),
        ChainSettings(levels=4, profile=0, base_pixels=48, last_pixels=6, max_lod0_delta_px=1.0,
                      transition=((0, 1), (1, 1)), search_orthographic=16, search_perspective=0,
                      search_seed=SEARCH_SEED, audit_orthographic=16, audit_perspective=0, audit_seed=AUDIT_SEED,
                      search_supersample=4, audit_supersample=8, max_supersample=16, candidate_budget=8,
                      beam_width=2),
        ChainOptions(action_trials=64, action_batch=32, ranking=4, overdraw_tiebreak=False))
    cases["ico2_edge_beam3"] = (
        meshes.icosphere(2, 0.03, 2),
        ChainSettings(levels=5, profile=0, base_pixels=40, last_pixels=5, max_lod0_delta_px=2.0,
                      transition=((0, 0.5), (1, 1.5)), search_orthographic=6, search_perspective=2,
                      search_seed=SEARCH_SEED, audit_orthographic=8, audit_perspective=2, audit_seed=AUDIT_SEED,
                      search_supersample=2, audit_supersample=4, max_supersample=8, candidate_budget=6,
                      beam_width=3, max_changed_area=0.5),
        ChainOptions(action_trials=24, action_batch=8, ranking=3, overdraw_tiebreak=False))
    cases["heightfield_plane"] = (
        meshes.jitter(meshes.heightfield(12, 0.25, 3), 0.005, 3, 0.5),
        ChainSettings(levels=3, profile=0, base_pixels=40, last_pixels=10, max_lod0_delta_px=1.5,
                      transition=((0, 1), (0.5, 1.25), (1, 1.5)), search_orthographic=6, search_perspective=2,
                      search_seed=SEARCH_SEED, audit_orthographic=10, audit_perspective=3, audit_seed=AUDIT_SEED,
                      search_supersample=2, audit_supersample=4, max_supersample=8, candidate_budget=5,
                      beam_width=2),
        ChainOptions(action_trials=40, action_batch=16, ranking=4, overdraw_tiebreak=False))
    cases["ico1_tiny"] = (
        meshes.icosphere(1, 0.02, 4),
        ChainSettings(levels=3, profile=0, base_pixels=24, last_pixels=6, max_lod0_delta_px=1.0,
                      transition=((0, 1), (1, 1)), search_orthographic=4, search_perspective=1,
                      search_seed=SEARCH_SEED, audit_orthographic=6, audit_perspective=1, audit_seed=AUDIT_SEED,
                      search_supersample=2, audit_supersample=4, max_supersample=8, candidate_budget=4,
                      beam_width=2),
        ChainOptions(action_trials=16, action_batch=8, ranking=4, overdraw_tiebreak=False))
    return cases
