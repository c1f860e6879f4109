"""ctypes binding of include/blitz_hip.h (libblitz_hip.so, built in-tree).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIBRARY_PATH = Path(__file__).resolve().parent / "libblitz_hip.so"

BLITZ_ABI_VERSION = 5
BLITZ_HIP_API_VERSION = 1
BLITZ_NEURAL_RASTER_HIP = 2

STATUS = {
    0: "BLITZ_OK",
    1: "BLITZ_INVALID_ARGUMENT",
    2: "BLITZ_OUT_OF_MEMORY",
    3: "BLITZ_INTERNAL_ERROR",
    4: "BLITZ_CANCELLED",
    5: "BLITZ_BUDGET_LIMITED",
    6: "BLITZ_UNAVAILABLE",
}


class BlitzStream(C.Structure):
    _fields_ = [("data", C.c_void_p), ("count", C.c_size_t), ("stride", C.c_size_t)]


class BlitzMesh(C.Structure):
    """blitz_mesh (reference include/blitz/blitz.h:28-37)."""
    _fields_ = [
        ("struct_size", C.c_uint32), ("abi_version", C.c_uint32),
        ("positions", BlitzStream), ("normals", BlitzStream), ("uv", BlitzStream),
        ("colors", BlitzStream), ("tangents", BlitzStream),
        ("indices", C.POINTER(C.c_uint32)), ("index_count", C.c_size_t),
        ("materials", C.POINTER(C.c_uint16)), ("material_count", C.c_size_t),
        ("double_sided", C.POINTER(C.c_uint8)), ("double_sided_count", C.c_size_t),
    ]


class Bounds(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("radius", C.c_double)]


class Camera(C.Structure):
    _fields_ = [
        ("right", C.c_float * 3), ("up", C.c_float * 3), ("forward", C.c_float * 3),
        ("distance", C.c_double), ("focal", C.c_double), ("scale", C.c_double),
        ("perspective", C.c_uint8),
    ]


CANCEL_FN = C.CFUNCTYPE(C.c_int, C.c_void_p)


class EvalSettingsC(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32), ("version", C.c_uint32),
        ("profile", C.c_uint8), ("supersample", C.c_uint8), ("max_supersample", C.c_uint8),
        ("force_two_sided", C.c_uint8), ("conservative_screen", C.c_uint8), ("reserved0", C.c_uint8 * 3),
        ("views_orthographic", C.c_uint16), ("views_perspective", C.c_uint16),
        ("views_rotation_seed", C.c_uint32),
        ("screen_size", C.c_double), ("limit", C.c_double), ("max_changed_area", C.c_double),
        ("normal_weight", C.c_double), ("color_weight", C.c_double), ("material_weight", C.c_double),
        ("cancelled", CANCEL_FN), ("user_data", C.c_void_p),
    ]


class MeasurementC(C.Structure):
    _fields_ = [
        ("error", C.c_double), ("coverage", C.c_double), ("coverage_upper", C.c_double),
        ("changed_area", C.c_double), ("normal_degrees", C.c_double),
        ("worst_view", C.c_uint32), ("changed_area_worst_view", C.c_uint32), ("views_evaluated", C.c_uint32),
        ("supersample", C.c_uint8), ("complete", C.c_uint8), ("passed", C.c_uint8),
        ("resource_limited", C.c_uint8), ("cancelled", C.c_uint8), ("reserved", C.c_uint8 * 3),
    ]


class AuditOptions(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32), ("version", C.c_uint32), ("device", C.c_int32),
        ("memory_mib", C.c_uint32), ("cache_rasters", C.c_uint8), ("face_split", C.c_uint8),
        ("autotune", C.c_uint8), ("reserved", C.c_uint8 * 5),
    ]


class ReduceReportC(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("passes", C.c_uint32), ("rounds", C.c_uint32),
                ("tail_collapses", C.c_uint32), ("collapses", C.c_uint64)]


class TrainSettingsC(C.Structure):
    _fields_ = [
        ("margin", C.c_float), ("penalty", C.c_float), ("lr", C.c_float), ("decay", C.c_float),
        ("max_norm", C.c_float), ("placement_weight", C.c_float), ("pool", C.c_uint32), ("precision", C.c_uint8),
    ]


class TrainResultC(C.Structure):
    _fields_ = [("loss", C.c_float), ("grad_norm", C.c_float), ("applied", C.c_uint8), ("step", C.c_uint32)]


class AuditTiming(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32), ("raster_tasks", C.c_uint32), ("distance_views", C.c_uint32),
        ("passes", C.c_uint32),
        ("raster_ms", C.c_double), ("compare_ms", C.c_double), ("distance_ms", C.c_double),
        ("total_ms", C.c_double), ("raster_triangles", C.c_uint64), ("mask_bytes", C.c_uint64),
        ("project_ms", C.c_double), ("band_ms", C.c_double), ("band_launches", C.c_uint32),
        ("patch_fallbacks", C.c_uint32), ("patch_ms", C.c_double), ("appearance_ms", C.c_double),
        ("attributed_views", C.c_uint32), ("split_launches", C.c_uint32), ("split_ms", C.c_double),
        ("raster_chunks", C.c_uint32), ("tuned_launches", C.c_uint32),
    ]


class ActionStats(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("states", C.c_uint32), ("ranked", C.c_uint64), ("trials", C.c_uint64),
                ("accepted", C.c_uint64), ("rejected", C.c_uint64), ("ms", C.c_double), ("build_ms", C.c_double),
                ("rank_ms", C.c_double), ("select_ms", C.c_double), ("trial_ms", C.c_double), ("gate_ms", C.c_double)]


class CurvePoint(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class BlitzSettings(C.Structure):
    """blitz_settings (reference include/blitz/blitz.h:41-63)."""
    _fields_ = [
        ("struct_size", C.c_uint32), ("abi_version", C.c_uint32),
        ("levels", C.c_uint8), ("profile", C.c_uint8), ("triangle_overhead_bps", C.c_uint16),
        ("max_added_vertex_bytes_bps", C.c_uint32),
        ("objective", C.c_uint8), ("beam_width", C.c_uint8), ("search_supersample", C.c_uint8),
        ("audit_supersample", C.c_uint8),
        ("max_supersample", C.c_uint8), ("prune", C.c_uint8), ("force_scalar", C.c_uint8),
        ("coupled_wedges", C.c_uint8),
        ("candidate_budget", C.c_uint16), ("search_ortho", C.c_uint16), ("search_perspective", C.c_uint16),
        ("audit_ortho", C.c_uint16), ("audit_perspective", C.c_uint16),
        ("search_seed", C.c_uint32), ("audit_seed", C.c_uint32),
        ("pixels_per_meter", C.c_double), ("meters_per_unit", C.c_double), ("base_pixels", C.c_double),
        ("last_pixels", C.c_double), ("max_lod0_delta_px", C.c_double),
        ("normal_weight", C.c_double), ("color_weight", C.c_double), ("material_weight", C.c_double),
        ("max_changed_area", C.c_double),
        ("transition", C.POINTER(CurvePoint)), ("transition_count", C.c_size_t),
        ("normal_importance", C.POINTER(CurvePoint)), ("normal_importance_count", C.c_size_t),
        ("attribute_importance", C.POINTER(CurvePoint)), ("attribute_importance_count", C.c_size_t),
        ("cancelled", CANCEL_FN), ("user_data", C.c_void_p),
    ]


class ChainOptionsC(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32), ("version", C.c_uint32), ("action_trials", C.c_uint32),
        ("action_batch", C.c_uint8), ("ranking", C.c_uint8), ("origin", C.c_uint8), ("output", C.c_uint8),
        ("confirmation", C.c_uint8), ("overdraw_tiebreak", C.c_uint8), ("reserved", C.c_uint8 * 6),
    ]


class ChainLodC(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32), ("indices", C.POINTER(C.c_uint32)), ("index_count", C.c_size_t),
        ("materials", C.POINTER(C.c_uint16)), ("material_count", C.c_size_t),
        ("screen_pixels", C.c_double), ("transition_limit", C.c_double), ("source_limit", C.c_double),
        ("adjacent", MeasurementC), ("source", MeasurementC),
    ]


class ChainInfoC(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32), ("levels", C.c_uint32), ("status", C.c_uint32), ("fallback_levels", C.c_uint32),
        ("candidate_evaluations", C.c_uint64), ("audit_evaluations", C.c_uint64 * 4),
        ("rejected_gates", C.c_uint64 * 4), ("area_rejected_gates", C.c_uint64 * 4),
        ("decoded", C.c_uint64), ("legal_collapses", C.c_uint64), ("rejected_collapses", C.c_uint64),
        ("action_ranked", C.c_uint64), ("action_trials", C.c_uint64), ("reference_rejections", C.c_uint64),
        ("bounded_audits", C.c_uint32), ("finalists", C.c_uint32), ("selected", C.c_uint32),
        ("proposals", C.c_uint32), ("gate_audits", C.c_uint64), ("confirm_audits", C.c_uint64),
        ("ms", C.c_double), ("executor_ms", C.c_double), ("gate_ms", C.c_double), ("offer_ms", C.c_double),
        ("confirm_ms", C.c_double), ("tiebreak_ms", C.c_double), ("build_ms", C.c_double), ("rank_ms", C.c_double),
        ("select_ms", C.c_double), ("trial_ms", C.c_double), ("speculative_ms", C.c_double),
        ("speculative_build_ms", C.c_double),
    ]


class ChainProposalC(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32), ("level", C.c_uint32), ("parent_triangles", C.c_uint32),
        ("start_triangles", C.c_uint32), ("target_triangles", C.c_uint32), ("final_triangles", C.c_uint32),
        ("trials", C.c_uint32), ("accepted_batches", C.c_uint32), ("accepted", C.c_uint64),
        ("rejected", C.c_uint64), ("screen_pixels", C.c_double), ("stop_reason", C.c_uint8), ("gate", C.c_uint8),
        ("reserved", C.c_uint8 * 6),
    ]


_P = C.c_void_p


class PlacementOptionsC(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("preserve_uv", C.c_int32), ("quantized", C.c_int32),
                ("low", C.c_float * 3), ("extent", C.c_float * 3), ("exact_bits", C.POINTER(C.c_uint32)),
                ("exact_words", C.c_size_t)]


class ModelInfoC(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("architecture", C.c_uint32), ("hidden_width", C.c_uint32),
                ("use", C.c_uint32), ("weight_count", C.c_uint64), ("weights", C.POINTER(C.c_float)),
                ("provenance_size", C.c_uint64), ("provenance", C.c_void_p)]
_ERR = [C.c_char_p, C.c_size_t]

# name -> (restype, argtypes). Every symbol include/blitz_hip.h declares.
PROTOTYPES = {
    "blitz_hip_api_version": (C.c_uint32, []),
    "blitz_hip_available": (C.c_int, [C.c_int32]),
    "blitz_hip_device_synchronize": (C.c_int, [C.c_int32]),
    "blitz_hip_audit_options_init": (C.c_int, [C.POINTER(AuditOptions), C.c_size_t]),
    "blitz_hip_eval_settings_init": (C.c_int, [C.POINTER(EvalSettingsC), C.c_size_t]),
    "blitz_hip_bounds_of": (C.c_int, [C.POINTER(BlitzMesh), C.POINTER(Bounds), *_ERR]),
    "blitz_hip_cameras": (C.c_int, [C.POINTER(Bounds), C.c_double, C.c_uint16, C.c_uint16, C.c_uint32,
                                    C.POINTER(Camera), C.c_size_t]),
    "blitz_hip_audit_create": (C.c_int, [C.POINTER(AuditOptions), C.POINTER(_P), *_ERR]),
    "blitz_hip_audit_destroy": (None, [_P]),
    "blitz_hip_mesh_upload": (C.c_int, [_P, C.POINTER(BlitzMesh), C.POINTER(_P), *_ERR]),
    "blitz_hip_mesh_release": (None, [_P]),
    "blitz_hip_audit_evaluate": (C.c_int, [_P, C.POINTER(BlitzMesh), C.POINTER(BlitzMesh), C.POINTER(Bounds),
                                           C.POINTER(EvalSettingsC), C.POINTER(MeasurementC), *_ERR]),
    "blitz_hip_audit_evaluate_device": (C.c_int, [_P, _P, _P, C.POINTER(Bounds), C.POINTER(EvalSettingsC),
                                                  C.POINTER(MeasurementC), *_ERR]),
    "blitz_hip_audit_evaluate_batch": (C.c_int, [_P, _P, C.POINTER(_P), C.c_size_t, C.POINTER(Bounds),
                                                 C.POINTER(EvalSettingsC), C.POINTER(MeasurementC), *_ERR]),
    "blitz_hip_audit_bind": (C.c_int, [_P, C.POINTER(BlitzMesh), C.POINTER(Bounds), C.POINTER(EvalSettingsC),
                                       C.POINTER(_P), *_ERR]),
    "blitz_hip_binding_release": (None, [_P]),
    "blitz_hip_audit_evaluate_patches": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                                   C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                                   C.POINTER(EvalSettingsC), C.POINTER(MeasurementC),
                                                   C.POINTER(C.c_uint32), *_ERR]),
    "blitz_hip_local_patches": (C.c_int, [C.POINTER(BlitzMesh), C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32,
                                          C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_size_t,
                                          C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_size_t,
                                          C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), *_ERR]),
    "blitz_hip_raster_coverage": (C.c_int, [_P, _P, C.POINTER(Bounds), C.POINTER(Camera), C.c_double, C.c_uint8,
                                            C.c_uint8, C.POINTER(C.c_uint64), C.c_size_t, C.POINTER(C.c_uint32),
                                            C.POINTER(C.c_uint8), *_ERR]),
    "blitz_hip_raster_pixels": (C.c_int, [_P, _P, C.POINTER(Bounds), C.POINTER(Camera), C.c_double, C.c_uint8,
                                          C.c_uint8, C.POINTER(C.c_uint8), C.POINTER(C.c_uint8),
                                          C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                          C.POINTER(C.c_uint16), C.c_size_t, C.POINTER(C.c_uint32),
                                          C.POINTER(C.c_uint8), *_ERR]),
    "blitz_hip_overlap": (C.c_int, [_P, _P, C.POINTER(Bounds), C.c_double, C.c_uint16, C.c_uint16, C.c_uint32,
                                    C.POINTER(C.c_double), *_ERR]),
    "blitz_hip_actions_create": (C.c_int, [C.c_int32, C.POINTER(BlitzMesh), C.POINTER(C.c_void_p), *_ERR]),
    "blitz_hip_actions_destroy": (None, [_P]),
    "blitz_hip_actions_info": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_double)]),
    "blitz_hip_actions_legal": (C.c_int, [_P, C.POINTER(C.c_uint32), C.c_size_t, *_ERR]),
    "blitz_hip_actions_independent": (C.c_int, [_P, C.POINTER(C.c_uint32), C.c_size_t, C.c_uint32, C.c_uint64,
                                                C.POINTER(C.c_uint32), C.POINTER(C.c_size_t), *_ERR]),
    "blitz_hip_actions_trial": (C.c_int, [_P, C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(C.c_uint32), C.c_size_t,
                                          C.POINTER(C.c_size_t), *_ERR]),
    "blitz_hip_actions_rank": (C.c_int, [_P, C.c_uint8, _P, _P, C.POINTER(C.c_uint32), C.c_size_t, *_ERR]),
    "blitz_hip_actions_vertex_features": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t, *_ERR]),
    "blitz_hip_policy_create": (C.c_int, [C.c_int32, C.c_uint32, C.c_uint32, C.POINTER(C.c_float), C.c_size_t,
                                          C.POINTER(_P), *_ERR]),
    "blitz_hip_policy_destroy": (None, [_P]),
    "blitz_hip_policy_predict": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t, C.POINTER(C.c_float), C.c_size_t,
                                           *_ERR]),
    "blitz_hip_placement_options_init": (C.c_int, [C.POINTER(PlacementOptionsC), C.c_size_t]),
    "blitz_hip_actions_placements": (C.c_int, [_P, _P, C.POINTER(C.c_float), C.c_uint64,
                                               C.POINTER(PlacementOptionsC), C.POINTER(C.c_float), C.c_size_t, *_ERR]),
    "blitz_hip_actions_mapping": (C.c_int, [_P, C.POINTER(C.c_uint32), C.c_size_t, *_ERR]),
    "blitz_hip_model_load": (C.c_int, [C.c_char_p, C.POINTER(_P), *_ERR]),
    "blitz_hip_model_destroy": (None, [_P]),
    "blitz_hip_model_sha256": (C.c_char_p, [_P]),
    "blitz_hip_model_info_of": (C.c_int, [_P, C.POINTER(ModelInfoC)]),
    "blitz_hip_model_save": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_float),
                                       C.c_size_t, C.c_char_p, C.c_size_t, *_ERR]),
    "blitz_hip_policy_from_model": (C.c_int, [C.c_int32, _P, C.POINTER(_P), *_ERR]),
    "blitz_hip_actions_rank_policy": (C.c_int, [_P, _P, C.POINTER(C.c_float), C.c_uint64, C.c_int32,
                                                C.POINTER(C.c_uint32), C.c_size_t, *_ERR]),
    "blitz_hip_actions_features": (C.c_int, [_P, C.POINTER(C.c_float), C.c_uint32, C.c_uint64, C.POINTER(C.c_float),
                                             C.c_size_t, *_ERR]),
    "blitz_hip_execute_actions": (C.c_int, [_P, C.POINTER(BlitzMesh), C.POINTER(Bounds), C.POINTER(EvalSettingsC),
                                            C.c_uint64, C.c_uint32, C.c_uint8, _P, C.c_uint8, C.POINTER(C.c_uint32),
                                            C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(ActionStats), *_ERR]),
    "blitz_hip_actions_patches": (C.c_int, [_P, C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(C.c_uint32),
                                            C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(C.c_uint32),
                                            C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(C.c_size_t),
                                            C.POINTER(C.c_size_t), *_ERR]),
    "blitz_hip_chain_options_init": (C.c_int, [C.POINTER(ChainOptionsC), C.c_size_t]),
    "blitz_hip_generate": (C.c_int, [_P, C.POINTER(BlitzMesh), C.POINTER(BlitzSettings), C.POINTER(ChainOptionsC),
                                     _P, C.POINTER(_P), *_ERR]),
    "blitz_hip_chain_destroy": (None, [_P]),
    "blitz_hip_schedule": (C.c_int, [C.POINTER(BlitzSettings), C.POINTER(Bounds), C.POINTER(C.c_double), C.c_size_t,
                                     *_ERR]),
    "blitz_hip_chain_info_of": (C.c_int, [_P, C.POINTER(ChainInfoC)]),
    "blitz_hip_chain_lod_of": (C.c_int, [_P, C.c_size_t, C.POINTER(ChainLodC)]),
    "blitz_hip_chain_finalist": (C.c_int, [_P, C.c_size_t, C.POINTER(C.c_uint32), C.c_size_t]),
    "blitz_hip_chain_proposal_of": (C.c_int, [_P, C.c_size_t, C.POINTER(ChainProposalC)]),
    "blitz_hip_coverage_distance": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_uint32,
                                              C.c_uint8, C.POINTER(C.c_double), *_ERR]),
    "blitz_hip_coverage_distance_noalloc": (C.c_int, [C.c_int32, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint8,
                                                      C.c_void_p, C.c_void_p, *_ERR]),
    "blitz_hip_audit_last_timing": (C.c_int, [_P, C.POINTER(AuditTiming)]),
    "blitz_hip_device_primitive": (C.c_int, [C.c_int32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int32,
                                             C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, *_ERR]),
    "blitz_hip_reduce_reuse": (C.c_int, [C.POINTER(BlitzMesh), C.c_size_t, C.POINTER(C.c_uint32), C.c_size_t,
                                         C.POINTER(C.c_size_t), *_ERR]),
    "blitz_hip_reduce_reuse_device": (C.c_int, [C.c_int32, C.POINTER(BlitzMesh), C.c_size_t, C.POINTER(C.c_uint32),
                                                C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(ReduceReportC), *_ERR]),
    "blitz_hip_scorer_weight_count": (C.c_size_t, []),
    "blitz_hip_scorer_create": (C.c_int, [C.c_int32, C.POINTER(C.c_float), C.c_size_t, C.POINTER(_P), *_ERR]),
    "blitz_hip_scorer_destroy": (None, [_P]),
    "blitz_hip_scorer_set_weights": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t, *_ERR]),
    "blitz_hip_graph_create": (C.c_int, [_P, C.POINTER(BlitzMesh), C.POINTER(Bounds), C.POINTER(_P), *_ERR]),
    "blitz_hip_graph_concat": (C.c_int, [_P, C.POINTER(_P), C.c_size_t, C.POINTER(_P), *_ERR]),
    "blitz_hip_graph_destroy": (None, [_P]),
    "blitz_hip_graph_info": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "blitz_hip_graph_features": (C.c_int, [_P, _P, C.POINTER(C.c_float), C.c_size_t]),
    "blitz_hip_graph_actions": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_size_t]),
    "blitz_hip_scorer_forward": (C.c_int, [_P, _P, C.c_uint8, C.POINTER(C.c_float), C.c_size_t, *_ERR]),
    "blitz_hip_scorer_rank": (C.c_int, [_P, _P, C.c_uint8, C.POINTER(C.c_uint32), C.c_size_t, *_ERR]),
    "blitz_hip_scorer_decode": (C.c_int, [_P, _P, C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(C.c_float), *_ERR]),
    "blitz_hip_scorer_teacher": (C.c_int, [_P, _P, C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(C.c_float),
                                           C.POINTER(C.c_float), *_ERR]),
    "blitz_hip_scorer_last_forward_ms": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "blitz_hip_train_settings_init": (None, [C.POINTER(TrainSettingsC)]),
    "blitz_hip_scorer_train_step": (C.c_int, [_P, _P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint8),
                                              C.POINTER(C.c_float), C.c_size_t, C.POINTER(TrainSettingsC),
                                              C.POINTER(TrainResultC), *_ERR]),
    "blitz_hip_scorer_gradients": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t]),
    "blitz_hip_scorer_weights": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t]),
    "blitz_hip_scorer_reset_optimizer": (C.c_int, [_P]),
    "blitz_hip_scorer_last_step_ms": (C.c_int, [_P, C.POINTER(C.c_double)]),
}

_lib: C.CDLL | None = None


def load() -> C.CDLL:
    """Loads libblitz_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIBRARY_PATH.exists():
        raise RuntimeError(f"{LIBRARY_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950) first")
    lib = C.CDLL(str(LIBRARY_PATH))
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class BlitzError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


def check(status: int, err: C.Array | None = None) -> None:
    if status != 0:
        raise BlitzError(status, err.value.decode(errors="replace") if err is not None else "")


def error_buffer() -> C.Array:
    return C.create_string_buffer(512)
