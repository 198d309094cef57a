"""Profile resolution and validation.

Mirrors the reference's resolve_profile (src/bmnd/profilevalidation.py:33-456):
the same checks, in the same order, with the same ValueError messages, and
the same in-place coercion of string fields to enums, and for the Gaussian
model the nf domain (psd.resolve_nf) with k = 0 on the conventional path.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .enums import (
    AggregationWeightDomain,
    AggregationWeightModel,
    AggregationWeightScope,
    BlockMatchDistance,
    NoiseModel,
    PoissonGroupMassConservation,
    PoissonHTMatchPolicy,
    PoissonVarianceSource,
    PoissonWienerGainMode,
    ReferenceScheduleMode,
    coerce_enum,
)
from .profiles import BMNDProfile
from .psd import resolve_nf
from .transforms import Transform, TransformMode, TransformType

_OPERATOR_TRANSFORMS = {
    TransformType.STARLET,
    TransformType.ATROUS,
    TransformType.B3SPLINE,
    TransformType.CURVELET,
}


@dataclass(frozen=True)
class ResolvedProfile:
    is_poisson: bool
    effective_nf: tuple[int, ...] | None
    conventional_nf: bool
    blockmatch_distance: BlockMatchDistance
    variance_k: int


def resolve_blockmatch_distance(distance, noise_model) -> BlockMatchDistance:
    """AUTO means Poisson deviance for Poisson noise and SSD otherwise
    (blockmatching.py:187-206)."""
    distance = coerce_enum(distance, BlockMatchDistance, "block-matching distance")
    noise_model = coerce_enum(noise_model, NoiseModel, "noise_model")
    if distance is BlockMatchDistance.AUTO:
        return (BlockMatchDistance.POISSON_DEVIANCE if noise_model is NoiseModel.POISSON
                else BlockMatchDistance.SSD)
    return distance


def _is_int(value) -> bool:
    return not isinstance(value, (bool, np.bool_)) and isinstance(value, (int, np.integer))


def _finite_nonnegative(name: str, value) -> float:
    msg = f"{name} must be finite and non-negative, got {value!r}."
    if isinstance(value, (bool, np.bool_)):
        raise ValueError(msg)
    try:
        v = float(value)
    except (TypeError, ValueError) as exc:
        raise ValueError(msg) from exc
    if not np.isfinite(v) or v < 0.0:
        raise ValueError(msg)
    return v


def _axis_values(name, values, shape, minimum, maximum=None) -> None:
    ndim = len(shape)
    try:
        vals = tuple(values)
    except TypeError as exc:
        raise ValueError(f"{name} must contain {ndim} integer values, got {values!r}.") from exc
    if len(vals) != ndim:
        raise ValueError(f"{name} must contain {ndim} values for shape {shape}, got {vals}.")
    if any(isinstance(v, bool) or not isinstance(v, (int, np.integer)) for v in vals):
        raise ValueError(f"{name} values must be integers, got {vals}.")
    if any(int(v) < minimum for v in vals):
        qualifier = "positive" if minimum == 1 else "non-negative"
        raise ValueError(f"{name} values must be {qualifier}, got {vals}.")
    if maximum is not None and any(int(v) > maximum[d] for d, v in enumerate(vals)):
        raise ValueError(f"{name} values cannot exceed volume shape {maximum}, got {vals}.")


def _stack_sizes(stage: str, minimum, maximum) -> None:
    for name, value in ((f"{stage}_min_stack_size", minimum), (f"{stage}_max_stack_size", maximum)):
        if not _is_int(value) or int(value) <= 0:
            raise ValueError(f"{name} must be a positive integer, got {value!r}.")
    if int(minimum) > int(maximum):
        raise ValueError(f"{stage}_min_stack_size cannot exceed {stage}_max_stack_size, "
                         f"got {minimum} > {maximum}.")


def _transform_list(name: str, transform) -> list[Transform]:
    if type(transform) is Transform:
        items = [transform]
    elif type(transform) is list:
        items = transform
    else:
        raise ValueError(f"{name} must be a Transform or list of Transform objects.")
    for t in items:
        if type(t) is not Transform:
            raise ValueError(f"{name} must contain only Transform objects.")
        if not isinstance(t.type, TransformType):
            raise ValueError(f"{name} contains an unknown transform type: {t.type!r}.")
        if not isinstance(t.mode, TransformMode):
            raise ValueError(f"{name} contains an unknown transform mode: {t.mode!r}.")
    return items


def resolve_profile(profile: BMNDProfile, volume_shape: tuple[int, ...]) -> ResolvedProfile:
    noise_model = coerce_enum(profile.noise_model, NoiseModel, "noise_model")
    profile.noise_model = noise_model
    mode = coerce_enum(profile.reference_schedule_mode, ReferenceScheduleMode,
                       "reference_schedule_mode")
    profile.reference_schedule_mode = mode
    if mode in {ReferenceScheduleMode.GENERATED, ReferenceScheduleMode.BALANCED,
                ReferenceScheduleMode.SPARSE}:
        _finite_nonnegative("reference_shift_density", profile.reference_shift_density)
        density = profile.reference_schedule_density
        if not _is_int(density) or int(density) <= 0:
            raise ValueError("reference_schedule_density must be a positive integer, "
                             f"got {density!r}.")

    for stage in ("ht", "wiener"):
        _axis_values(f"{stage}_block_size", getattr(profile, f"{stage}_block_size"),
                     volume_shape, 1, volume_shape)
        _axis_values(f"{stage}_step", getattr(profile, f"{stage}_step"), volume_shape, 1)
        _axis_values(f"{stage}_search_window", getattr(profile, f"{stage}_search_window"),
                     volume_shape, 0)
        _stack_sizes(stage, getattr(profile, f"{stage}_min_stack_size"),
                     getattr(profile, f"{stage}_max_stack_size"))
        _transform_list(f"{stage}_transform", getattr(profile, f"{stage}_transform"))
        _finite_nonnegative(f"{stage}_match_threshold", getattr(profile, f"{stage}_match_threshold"))
        _finite_nonnegative(f"{stage}_kaiser_beta", getattr(profile, f"{stage}_kaiser_beta"))

    _finite_nonnegative("ht_lambda_threshold", profile.ht_lambda_threshold)
    _finite_nonnegative("wiener_variance_scale", profile.wiener_variance_scale)
    if not _is_int(profile.ref_batch_size) or int(profile.ref_batch_size) <= 0:
        raise ValueError(f"ref_batch_size must be a positive integer, got {profile.ref_batch_size!r}.")

    configured = coerce_enum(profile.blockmatch_distance, BlockMatchDistance, "blockmatch_distance")
    profile.blockmatch_distance = configured
    distance = resolve_blockmatch_distance(configured, noise_model)
    is_poisson = noise_model is NoiseModel.POISSON

    if not _is_int(profile.k) or int(profile.k) < 0:
        raise ValueError(f"k must be a non-negative integer, got {profile.k!r}.")
    variance_k = int(profile.k)

    for stage in ("ht", "wiener"):
        names = (f"{stage}_weight_model", f"{stage}_weight_domain", f"{stage}_weight_scope")
        model = coerce_enum(getattr(profile, names[0]), AggregationWeightModel, names[0])
        domain = coerce_enum(getattr(profile, names[1]), AggregationWeightDomain, names[1])
        scope = coerce_enum(getattr(profile, names[2]), AggregationWeightScope, names[2])
        setattr(profile, names[0], model)
        setattr(profile, names[1], domain)
        setattr(profile, names[2], scope)
        valid = {AggregationWeightModel.CLASSIC, AggregationWeightModel.VARIANCE}
        if stage == "wiener":
            valid.add(AggregationWeightModel.RISK)
        if model not in valid:
            expected = "', '".join(sorted(valid))
            raise ValueError(f"Unknown {names[0]}: {model!r} (expected '{expected}').")
        transforms = _transform_list(f"{stage}_transform", getattr(profile, f"{stage}_transform"))
        all_synthesis = domain is AggregationWeightDomain.WINDOWED_SYNTHESIS
        group_synthesis = scope is AggregationWeightScope.PATCH
        for t in transforms:
            on_group = t.mode in {TransformMode.ND, TransformMode.GROUP}
            if not all_synthesis and not (group_synthesis and on_group):
                continue
            if t.type in _OPERATOR_TRANSFORMS:
                raise ValueError(f"{stage.upper()} {domain}/{scope} weighting requires "
                                 "matrix-based transforms.")
            if t.type is TransformType.FFT:
                raise ValueError(f"{stage.upper()} {domain}/{scope} weighting requires "
                                 "real matrix transforms.")

    if profile.sharpen_alpha != 1.0 or profile.sharpen_alpha_3d != 1.0:
        raise ValueError("Aggregation weighting requires sharpening exponents of 1.")
    gamma = float(profile.gamma)
    if not np.isfinite(gamma) or gamma < 0.0:
        raise ValueError(f"gamma must be finite and non-negative, got {profile.gamma}.")

    if is_poisson:
        for name, enum_type in (("poisson_ht_match_policy", PoissonHTMatchPolicy),
                                ("poisson_wiener_gain_mode", PoissonWienerGainMode),
                                ("poisson_group_mass_conservation", PoissonGroupMassConservation)):
            setattr(profile, name, coerce_enum(getattr(profile, name), enum_type, name))
        for name in ("poisson_variance_source_wiener", "poisson_wiener_match_source"):
            setattr(profile, name, coerce_enum(getattr(profile, name), PoissonVarianceSource, name))
        scale = float(profile.poisson_count_scale)
        if not np.isfinite(scale) or scale <= 0.0:
            raise ValueError("poisson_count_scale must be finite and positive, "
                             f"got {profile.poisson_count_scale}.")
        _finite_nonnegative("poisson_variance_floor", profile.poisson_variance_floor)
        mass = profile.poisson_group_mass_conservation
        if mass in {PoissonGroupMassConservation.WIENER, PoissonGroupMassConservation.BOTH}:
            if profile.sharpen_alpha != 1.0 or profile.sharpen_alpha_3d != 1.0:
                raise ValueError("Wiener group-mass conservation requires sharpening exponents of 1.")
            if profile.wiener_weight_model in {AggregationWeightModel.VARIANCE,
                                               AggregationWeightModel.RISK} and any(
                    t.type is TransformType.FFT
                    for t in _transform_list("wiener_transform", profile.wiener_transform)):
                raise ValueError("Mass-projected Poisson Wiener weighting requires real matrix transforms.")
        for stage, enabled in (
                ("ht", mass in {PoissonGroupMassConservation.HT, PoissonGroupMassConservation.BOTH}),
                ("wiener", mass in {PoissonGroupMassConservation.WIENER,
                                    PoissonGroupMassConservation.BOTH})):
            if not enabled:
                continue
            transforms = _transform_list(f"{stage}_transform", getattr(profile, f"{stage}_transform"))
            if any(t.type in _OPERATOR_TRANSFORMS for t in transforms):
                raise ValueError(f"{stage.upper()} group-mass conservation requires matrix-based transforms.")
            projected = getattr(profile, f"{stage}_weight_model") in {
                AggregationWeightModel.VARIANCE, AggregationWeightModel.RISK}
            if projected and any(t.type is TransformType.FFT for t in transforms):
                raise ValueError(f"Mass-projected Poisson {stage.upper()} weighting requires "
                                 "real matrix transforms.")
        max_count = profile.poisson_ht_match_moment_max_count
        if not _is_int(max_count) or int(max_count) <= 0:
            raise ValueError("poisson_ht_match_moment_max_count must be a positive integer, "
                             f"got {max_count!r}.")
        policy = profile.poisson_ht_match_policy
        if policy is not PoissonHTMatchPolicy.FIXED and distance not in {
                BlockMatchDistance.POISSON_DEVIANCE, BlockMatchDistance.PEARSON,
                BlockMatchDistance.ANSCOMBE_SSD}:
            raise ValueError(f"Poisson HT match policy {policy!r} requires 'poisson_deviance', "
                             f"'pearson', or 'anscombe_ssd', got {distance!r}.")
        if policy is PoissonHTMatchPolicy.REFERENCE_FINITE_COUNT:
            for name, value, positive in (
                    ("poisson_ht_match_structure_beta", float(profile.poisson_ht_match_structure_beta), False),
                    ("poisson_ht_match_intensity_power", profile.poisson_ht_match_intensity_power, True)):
                if not np.isfinite(value) or value < 0.0 or (positive and value == 0.0):
                    qualifier = "positive" if positive else "non-negative"
                    raise ValueError(f"{name} must be finite and {qualifier}, got {value}.")
    effective_nf, conventional_nf = None, False
    if not is_poisson:
        # profilevalidation.py:445-448
        effective_nf, conventional_nf = resolve_nf(profile.nf, volume_shape)
        if conventional_nf:
            variance_k = 0
    return ResolvedProfile(is_poisson=is_poisson, effective_nf=effective_nf,
                           conventional_nf=conventional_nf, blockmatch_distance=distance,
                           variance_k=variance_k)
