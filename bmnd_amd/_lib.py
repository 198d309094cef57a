"""ctypes binding of libbmnd_hip.so (declarations in include/bmnd_hip.h).

torch is imported first so that the process has exactly one HIP runtime:
the library's libamdhip64.so.7 dependency resolves to the one torch already
loaded.  There is no fallback: if the library is missing the import fails
loudly.
"""

from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (must precede the library load)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbmnd_hip.so")
_LIB: ctypes.CDLL | None = None

EXPORTS = (
    "bmnd_hip_version",
    "bmnd_hip_last_error",
    "bmnd_hip_configs",
    "bmnd_hip_variance_volume",
    "bmnd_hip_count_range",
    "bmnd_hip_blockmatch",
    "bmnd_hip_blockmatch_thresholds",
    "bmnd_hip_blockmatch_standardized",
    "bmnd_hip_blockmatch_ssd",
    "bmnd_hip_blockmatch_ssd_poisson",
    "bmnd_hip_value_range",
    "bmnd_hip_gaussian_variance",
    "bmnd_hip_median_mad",
    "bmnd_hip_log2_check",
    "bmnd_hip_matrix_table_size",
    "bmnd_hip_exact_std",
    "bmnd_hip_exact_std_pairs",
    "bmnd_hip_filter_groups",
    "bmnd_hip_filter_groups_sigma",
    "bmnd_hip_exact_risk_workspace",
    "bmnd_hip_filter_groups_exact",
    "bmnd_hip_reference_map",
    "bmnd_hip_aggregate",
    "bmnd_hip_aggregate_sigma",
    "bmnd_hip_aggregate_norm",
    "bmnd_hip_mass_correct",
    "bmnd_hip_poisson_global_sigma",
    "bmnd_hip_np_sum_host",
    "bmnd_hip_aggregate_tile",
    "bmnd_hip_aggregate_force_tile",
    "bmnd_hip_blockmatch_batch",
    "bmnd_hip_exact_std_batch",
    "bmnd_hip_exact_std_pairs_batch",
    "bmnd_hip_filter_groups_batch",
    "bmnd_hip_reference_map_batch",
    "bmnd_hip_aggregate_batch",
    "bmnd_hip_stage_workspace",
    "bmnd_hip_stage",
    "bmnd_hip_init",
    "bmnd_hip_stage_noalloc",
    "bmnd_hip_pipeline_create",
    "bmnd_hip_pipeline_create_many",
    "bmnd_hip_pipeline_launch",
    "bmnd_hip_pipeline_status",
    "bmnd_hip_pipeline_destroy",
    "bmnd_hip_schedule_capacity",
    "bmnd_hip_schedule_generated",
    "bmnd_hip_schedule_axis_shifts",
    # the pilot network (include/bmnd_conv.h)
    "bmnd_conv_kpad",
    "bmnd_conv3d_forward",
    "bmnd_conv3d_wgrad_parts",
    "bmnd_conv3d_wgrad",
    "bmnd_conv_in_forward",
    "bmnd_conv_out_forward",
    "bmnd_conv_out_backward",
    "bmnd_conv_small_wgrad",
    "bmnd_conv_bias_grad",
    "bmnd_conv_pool2",
    "bmnd_conv_skip_grad",
    "bmnd_conv_up_grad",
    "bmnd_conv_mask_grad",
    "bmnd_conv_flip_weights",
    "bmnd_conv_mse",
    "bmnd_adam_step",
    "bmnd_to_bf16",
    "bmnd_anscombe",
    "bmnd_anscombe_moments",
    "bmnd_pad_edge",
    "bmnd_compare_crops",
)


class HipError(RuntimeError):
    pass


class PipelineDesc(ctypes.Structure):
    """bmnd_hip_pipeline_desc (include/bmnd_hip.h)."""
    _fields_ = [
        ("shape", ctypes.c_int64 * 3),
        ("block", (ctypes.c_int32 * 3) * 2),
        ("radius", (ctypes.c_int32 * 3) * 2),
        ("ref_lin", ctypes.c_void_p * 2),
        ("nrefs", ctypes.c_int64 * 2),
        ("threshold", ctypes.c_float * 2),
        ("max_matches", ctypes.c_int32 * 2),
        ("min_matches", ctypes.c_int32 * 2),
        ("mode", ctypes.c_int32 * 2),
        ("table_max", ctypes.c_int32),
        ("matrices", ctypes.c_void_p * 2),
        ("kaiser", ctypes.c_void_p * 2),
        ("shrink", ctypes.c_float * 2),
        ("k", ctypes.c_int32),
        ("wiener_gain", ctypes.c_int32),
        ("count_scale", ctypes.c_double),
        ("variance_floor", ctypes.c_double),
        ("match_observation", ctypes.c_int32),
        ("variance_observation", ctypes.c_int32),
        ("input", ctypes.c_void_p),
        ("pilot", ctypes.c_void_p),
        ("output", ctypes.c_void_p),
        ("sel", ctypes.c_void_p * 2),
        ("counts", ctypes.c_void_p * 2),
        ("accepted", ctypes.c_void_p * 2),
        ("ninputs", ctypes.c_int32),
        ("batch", ctypes.c_void_p * 2),
        ("batch_d", ctypes.c_void_p * 2),
        ("ref_input", ctypes.c_void_p * 2),
    ]


def _declare(L: ctypes.CDLL) -> None:
    vp, i32, i64, f32, f64 = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64,
                              ctypes.c_float, ctypes.c_double)
    L.bmnd_hip_version.restype = i32
    L.bmnd_hip_last_error.restype = ctypes.c_char_p
    L.bmnd_hip_configs.restype = ctypes.c_char_p
    L.bmnd_hip_variance_volume.argtypes = [vp, vp, i64, f64, f64, vp]
    L.bmnd_hip_variance_volume.restype = i32
    L.bmnd_hip_count_range.argtypes = [vp, i64, vp, vp]
    L.bmnd_hip_count_range.restype = i32
    L.bmnd_hip_blockmatch.argtypes = [vp, vp, vp, vp, vp, i64, f32, i32, i32, i32, i32, f64,
                                      i32, i32, vp, vp, vp, vp]
    L.bmnd_hip_blockmatch.restype = i32
    L.bmnd_hip_blockmatch_thresholds.argtypes = [vp, vp, vp, vp, vp, vp, i64, i32, i32, i32, i32,
                                                 f64, i32, i32, vp, vp, vp, vp]
    L.bmnd_hip_blockmatch_thresholds.restype = i32
    L.bmnd_hip_blockmatch_standardized.argtypes = [vp, vp, vp, vp, vp, i64, f32, i32, i32, i32,
                                                   i32, f64, i32, vp, i32, vp, vp, vp, vp]
    L.bmnd_hip_blockmatch_standardized.restype = i32
    L.bmnd_hip_log2_check.argtypes = [vp]
    L.bmnd_hip_log2_check.restype = i32
    L.bmnd_hip_matrix_table_size.argtypes = [i32, vp]
    L.bmnd_hip_matrix_table_size.restype = i64
    L.bmnd_hip_exact_std.argtypes = [vp, vp, vp, vp, vp, vp, i64, i32, vp, i32, vp, vp]
    L.bmnd_hip_exact_std.restype = i32
    L.bmnd_hip_exact_std_pairs.argtypes = L.bmnd_hip_exact_std.argtypes
    L.bmnd_hip_exact_std_pairs.restype = i32
    L.bmnd_hip_blockmatch_ssd.argtypes = [vp, vp, vp, vp, vp, i64, f32, i32, i32, i32, vp, f32,
                                          vp, vp, vp, vp]
    L.bmnd_hip_blockmatch_ssd.restype = i32
    L.bmnd_hip_blockmatch_ssd_poisson.argtypes = [vp, vp, vp, vp, vp, i64, f32, i32, i32, i32,
                                                  f32, f64, vp, vp, vp, vp]
    L.bmnd_hip_blockmatch_ssd_poisson.restype = i32
    L.bmnd_hip_value_range.argtypes = [vp, i64, vp, vp]
    L.bmnd_hip_value_range.restype = i32
    L.bmnd_hip_gaussian_variance.argtypes = [vp, vp, vp, vp, vp, vp, vp, i64, i32, vp, i32, vp,
                                             vp]
    L.bmnd_hip_gaussian_variance.restype = i32
    L.bmnd_hip_median_mad.argtypes = [vp, i64, vp, vp, vp]
    L.bmnd_hip_median_mad.restype = i32
    L.bmnd_hip_filter_groups.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, vp, i64, i32, vp, vp,
                                         f32, i32, f64, vp, vp, vp]
    L.bmnd_hip_filter_groups.restype = i32
    L.bmnd_hip_reference_map.argtypes = [vp, i64, vp, i64, vp]
    L.bmnd_hip_reference_map.restype = i32
    L.bmnd_hip_aggregate.argtypes = [vp, vp, vp, vp, vp, i64, i32, vp, vp, vp, vp, vp, vp, vp]
    L.bmnd_hip_aggregate.restype = i32
    L.bmnd_hip_filter_groups_sigma.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, vp, i64, i32, vp,
                                               vp, f32, i32, f64, vp, vp, vp, vp]
    L.bmnd_hip_filter_groups_sigma.restype = i32
    L.bmnd_hip_exact_risk_workspace.argtypes = [i64, i32, i32, vp]
    L.bmnd_hip_exact_risk_workspace.restype = i64
    L.bmnd_hip_filter_groups_exact.argtypes = L.bmnd_hip_filter_groups_sigma.argtypes[:-1] + [
        vp, i64, vp]
    L.bmnd_hip_filter_groups_exact.restype = i32
    L.bmnd_hip_filter_groups_sigma.restype = i32
    L.bmnd_hip_aggregate_sigma.argtypes = [vp, vp, vp, vp, vp, i64, i32, vp, vp, vp, vp, vp, vp,
                                           vp, i32, f32, vp, vp]
    L.bmnd_hip_aggregate_sigma.restype = i32
    L.bmnd_hip_aggregate_norm.argtypes = [i32, vp, vp, vp, vp, vp, i64, i32, vp, vp, vp, vp, vp,
                                          vp, vp, i32, f32, vp, vp]
    L.bmnd_hip_aggregate_norm.restype = i32
    L.bmnd_hip_mass_correct.argtypes = [vp, vp, vp, vp, i64, i32, vp, vp, vp, vp, vp, vp, vp, vp]
    L.bmnd_hip_mass_correct.restype = i32
    L.bmnd_hip_poisson_global_sigma.argtypes = [vp, i64, f64, ctypes.POINTER(ctypes.c_double), vp]
    L.bmnd_hip_poisson_global_sigma.restype = i32
    L.bmnd_hip_np_sum_host.argtypes = [vp, i64]
    L.bmnd_hip_np_sum_host.restype = f32
    L.bmnd_hip_aggregate_tile.argtypes = [vp, vp]
    L.bmnd_hip_aggregate_tile.restype = i32
    L.bmnd_hip_aggregate_force_tile.argtypes = [i32]
    L.bmnd_hip_aggregate_force_tile.restype = i32
    L.bmnd_hip_blockmatch_batch.argtypes = [vp, vp, vp, i32, vp, vp, vp, vp, vp, i64, f32, i32,
                                            i32, i32, i32, f64, i32, i32, vp, i32, vp, vp, vp,
                                            vp]
    L.bmnd_hip_blockmatch_batch.restype = i32
    L.bmnd_hip_exact_std_batch.argtypes = [vp, vp, vp, i32, vp, vp, vp, vp, vp, i64, i32, vp,
                                           i32, vp, vp]
    L.bmnd_hip_exact_std_batch.restype = i32
    L.bmnd_hip_exact_std_pairs_batch.argtypes = L.bmnd_hip_exact_std_batch.argtypes
    L.bmnd_hip_exact_std_pairs_batch.restype = i32
    L.bmnd_hip_filter_groups_batch.argtypes = [i32, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp, vp,
                                               i64, i32, vp, vp, f32, i32, f64, vp, vp, vp]
    L.bmnd_hip_filter_groups_batch.restype = i32
    L.bmnd_hip_reference_map_batch.argtypes = [vp, vp, i64, vp, vp, i64, vp]
    L.bmnd_hip_reference_map_batch.restype = i32
    L.bmnd_hip_aggregate_batch.argtypes = [vp, vp, vp, vp, vp, i64, i32, vp, vp, i32, vp, vp, vp,
                                           vp, vp, vp]
    L.bmnd_hip_aggregate_batch.restype = i32
    L.bmnd_hip_stage_workspace.argtypes = [vp, vp, i64, i32, i32]
    L.bmnd_hip_stage_workspace.restype = i64
    L.bmnd_hip_stage.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, vp, i64, f32, i32, i32, f64, i32,
                                 vp, vp, f32, i32, vp, i64, vp, vp, vp, vp, vp]
    L.bmnd_hip_stage.restype = i32
    L.bmnd_hip_init.argtypes = []
    L.bmnd_hip_init.restype = i32
    L.bmnd_hip_stage_noalloc.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, vp, i64, f32, i32, i32,
                                         f64, i32, vp, vp, f32, i32, vp, vp, vp, vp, vp]
    L.bmnd_hip_stage_noalloc.restype = i32
    L.bmnd_hip_pipeline_create.argtypes = [ctypes.POINTER(PipelineDesc),
                                           ctypes.POINTER(ctypes.c_void_p)]
    L.bmnd_hip_pipeline_create.restype = i32
    L.bmnd_hip_pipeline_create_many.argtypes = [ctypes.POINTER(PipelineDesc), i32,
                                                ctypes.POINTER(vp)]
    L.bmnd_hip_pipeline_create_many.restype = i32
    L.bmnd_hip_pipeline_launch.argtypes = [vp, vp]
    L.bmnd_hip_pipeline_launch.restype = i32
    L.bmnd_hip_pipeline_status.argtypes = [vp, vp]
    L.bmnd_hip_pipeline_status.restype = i32
    L.bmnd_hip_pipeline_destroy.argtypes = [vp]
    L.bmnd_hip_pipeline_destroy.restype = None
    L.bmnd_hip_schedule_capacity.argtypes = [i32, vp, vp, vp, f64, i32]
    L.bmnd_hip_schedule_capacity.restype = i64
    L.bmnd_hip_schedule_generated.argtypes = [i32, vp, vp, vp, f64, i32, vp, vp, i64, vp, vp]
    L.bmnd_hip_schedule_generated.restype = i32
    L.bmnd_hip_schedule_axis_shifts.argtypes = [i32, i32, i32, f64, vp]
    L.bmnd_hip_schedule_axis_shifts.restype = i32
    # the pilot network
    L.bmnd_conv_kpad.argtypes = [i32]
    L.bmnd_conv_kpad.restype = i64
    L.bmnd_conv3d_forward.argtypes = [vp, vp, i32, i32, i32, vp, vp, i32, vp, i32, i32, i32, i32,
                                      i32, i32, vp, vp]
    L.bmnd_conv3d_wgrad_parts.argtypes = [i32, i32, i32, i32, i32]
    L.bmnd_conv3d_wgrad_parts.restype = i64
    L.bmnd_conv3d_wgrad.argtypes = [vp, vp, i32, i32, i32, vp, i32, vp, i32, vp, i32, i32, i32,
                                    i32, vp]
    L.bmnd_conv_in_forward.argtypes = [vp, vp, vp, vp, i32, i32, i32, i32, vp]
    L.bmnd_conv_out_forward.argtypes = [vp, vp, vp, vp, vp, i32, i32, i32, i32, vp]
    L.bmnd_conv_out_backward.argtypes = [vp, vp, vp, vp, i32, i32, i32, i32, vp]
    L.bmnd_conv_small_wgrad.argtypes = [i32, vp, vp, vp, vp, vp, i32, vp, i32, i32, i32, i32, vp]
    L.bmnd_conv_bias_grad.argtypes = [vp, vp, i32, vp, i32, vp, i64, vp]
    L.bmnd_conv_pool2.argtypes = [vp, vp, i32, i32, i32, i32, i32, vp]
    L.bmnd_conv_skip_grad.argtypes = [vp, i32, i32, vp, vp, vp, i32, i32, i32, i32, i32, vp]
    L.bmnd_conv_up_grad.argtypes = [vp, i32, vp, vp, i32, i32, i32, i32, i32, vp]
    L.bmnd_conv_mask_grad.argtypes = [vp, vp, vp, i64, vp]
    L.bmnd_conv_flip_weights.argtypes = [vp, i32, i32, vp, vp]
    L.bmnd_conv_mse.argtypes = [vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, f32, vp, vp]
    L.bmnd_adam_step.argtypes = [vp, vp, vp, vp, vp, i64, f32, f32, f32, f32, i32, vp]
    L.bmnd_to_bf16.argtypes = [vp, vp, i64, vp]
    L.bmnd_anscombe.argtypes = [vp, vp, i64, i32, f32, f32, vp]
    L.bmnd_anscombe_moments.argtypes = [vp, i64, vp, vp, vp]
    L.bmnd_anscombe_moments.restype = i32
    L.bmnd_pad_edge.argtypes = [vp, vp, i32, i32, i32, i32, i32, i32, i32, vp]
    L.bmnd_compare_crops.argtypes = [vp, vp, vp, i32, i32, i32, i32, i32, i32, i32, i32, vp, i32,
                                     vp]
    for name in ("bmnd_conv3d_forward", "bmnd_conv3d_wgrad", "bmnd_conv_in_forward",
                 "bmnd_conv_out_forward", "bmnd_conv_out_backward", "bmnd_conv_small_wgrad",
                 "bmnd_conv_bias_grad", "bmnd_conv_pool2", "bmnd_conv_skip_grad",
                 "bmnd_conv_up_grad", "bmnd_conv_mask_grad", "bmnd_conv_flip_weights",
                 "bmnd_conv_mse", "bmnd_adam_step", "bmnd_to_bf16", "bmnd_anscombe",
                 "bmnd_pad_edge", "bmnd_compare_crops"):
        getattr(L, name).restype = i32


def lib() -> ctypes.CDLL:
    """Load the HIP library (once).  Raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise HipError(
                f"{LIB_PATH} is missing: the MI355X kernels are not built. "
                "Run `python -c 'import __graft_entry__ as g; g.build()'` first.")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _LIB = L
    return _LIB


def check(status: int) -> None:
    if status != 0:
        msg = lib().bmnd_hip_last_error().decode(errors="replace")
        raise HipError(f"libbmnd_hip call failed ({status}): {msg}")
