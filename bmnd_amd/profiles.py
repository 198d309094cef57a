"""Profile dataclasses of the bmnd API.

Field names, types and defaults follow the reference's src/bmnd/profiles.py
(BMNDProfile :18-174, BM2DProfile :177-222, BM3DProfile :225-288,
BM4DProfile :291-352, BM5DProfile :355-405) so existing user code keeps
working; the MI355X path executes the Poisson branch (noise_model=POISSON).
"""

from __future__ import annotations

from dataclasses import dataclass, field

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
)
from .transforms import Transform, TransformMode, TransformType


def _haar_bior() -> list[Transform]:
    return [Transform(TransformType.HAAR, TransformMode.GROUP),
            Transform(TransformType.BIOR15, TransformMode.SPATIAL)]


def _haar_dct() -> list[Transform]:
    return [Transform(TransformType.HAAR, TransformMode.GROUP),
            Transform(TransformType.DCT, TransformMode.SPATIAL)]


@dataclass
class BMNDProfile:
    """All algorithm hyper-parameters of bmnd()."""

    noise_model: NoiseModel = NoiseModel.GAUSSIAN
    # reference schedule
    reference_schedule_mode: ReferenceScheduleMode = ReferenceScheduleMode.GENERATED
    reference_shift_density: float = 2.0
    reference_schedule_density: int = 2
    # Poisson
    poisson_variance_source_wiener: PoissonVarianceSource = PoissonVarianceSource.PILOT
    poisson_variance_floor: float = 1e-10
    poisson_count_scale: float = 1.0
    poisson_ht_match_policy: PoissonHTMatchPolicy = PoissonHTMatchPolicy.FIXED
    poisson_wiener_match_source: PoissonVarianceSource = PoissonVarianceSource.PILOT
    poisson_ht_match_structure_beta: float = 0.1
    poisson_ht_match_intensity_power: float = 1.0
    poisson_ht_match_moment_max_count: int = 512
    poisson_wiener_gain_mode: PoissonWienerGainMode = PoissonWienerGainMode.VARIANCE_SCALED
    poisson_group_mass_conservation: PoissonGroupMassConservation = (
        PoissonGroupMassConservation.NONE
    )
    # noise variance and covariance
    nf: tuple[int, ...] | int | None = field(default_factory=tuple)
    k: int = 4
    # aggregation weighting
    ht_weight_model: AggregationWeightModel = AggregationWeightModel.VARIANCE
    ht_weight_domain: AggregationWeightDomain = AggregationWeightDomain.COEFFICIENT
    ht_weight_scope: AggregationWeightScope = AggregationWeightScope.PATCH
    wiener_weight_model: AggregationWeightModel = AggregationWeightModel.VARIANCE
    wiener_weight_domain: AggregationWeightDomain = AggregationWeightDomain.COEFFICIENT
    wiener_weight_scope: AggregationWeightScope = AggregationWeightScope.PATCH
    # refiltering
    denoise_residual: bool = False
    residual_threshold: float = 3.0
    max_pad_size: int = 0
    # block matching
    gamma: float = 1.0
    blockmatch_distance: BlockMatchDistance = BlockMatchDistance.AUTO
    # hard thresholding
    ht_use_soft_thresholding: bool = False
    ht_block_size: tuple[int, ...] = field(default_factory=tuple)
    ht_step: tuple[int, ...] = field(default_factory=tuple)
    ht_search_window: tuple[int, ...] = field(default_factory=tuple)
    ht_max_stack_size: int = field(default_factory=int)
    ht_min_stack_size: int = 2
    ht_match_threshold: float = 2.9527
    ht_lambda_threshold: float = 1.0
    ht_kaiser_beta: float = 2.0
    # Wiener filtering
    wiener_block_size: tuple[int, ...] = field(default_factory=tuple)
    wiener_step: tuple[int, ...] = field(default_factory=tuple)
    wiener_search_window: tuple[int, ...] = field(default_factory=tuple)
    wiener_max_stack_size: int = field(default_factory=int)
    wiener_min_stack_size: int = 2
    wiener_match_threshold: float = 0.7689
    wiener_variance_scale: float = 1.0
    wiener_kaiser_beta: float = 2.0
    # transforms
    ht_transform: Transform | list[Transform] = field(
        default_factory=lambda: Transform(TransformType.BIOR15, TransformMode.ND))
    wiener_transform: Transform | list[Transform] = field(
        default_factory=lambda: Transform(TransformType.DCT, TransformMode.ND))
    # other
    sharpen_alpha: float = 1.0
    sharpen_alpha_3d: float = 1.0
    ref_batch_size: int = 64


@dataclass
class BM2DProfile(BMNDProfile):
    """One-dimensional BMND template."""

    noise_model: NoiseModel = NoiseModel.GAUSSIAN
    gamma: float = 3.0
    nf: tuple[int] | int | None = (32,)
    k: int = 4
    ht_use_soft_thresholding: bool = False
    ht_block_size: tuple[int] = (64,)
    ht_step: tuple[int] = (16,)
    ht_search_window: tuple[int] = (128,)
    ht_max_stack_size: int = 16
    ht_min_stack_size: int = 2
    ht_match_threshold: float = 2.9527
    ht_lambda_threshold: float = 3.0
    ht_kaiser_beta: float = 2.0
    wiener_block_size: tuple[int] = (64,)
    wiener_step: tuple[int] = (16,)
    wiener_search_window: tuple[int] = (128,)
    wiener_max_stack_size: int = 32
    wiener_min_stack_size: int = 2
    wiener_match_threshold: float = 0.3937
    wiener_variance_scale: float = 0.4
    wiener_kaiser_beta: float = 2.0
    ht_transform: Transform | list[Transform] = field(default_factory=_haar_bior)
    wiener_transform: Transform | list[Transform] = field(default_factory=_haar_dct)


@dataclass
class BM3DProfile(BMNDProfile):
    """Official BM3D compatibility values (2-D images)."""

    noise_model: NoiseModel = NoiseModel.GAUSSIAN
    gamma: float = 3.0
    nf: tuple[int, int] | int | None = (32, 32)
    k: int = 4
    ht_weight_model: AggregationWeightModel = AggregationWeightModel.VARIANCE
    ht_weight_domain: AggregationWeightDomain = AggregationWeightDomain.COEFFICIENT
    ht_weight_scope: AggregationWeightScope = AggregationWeightScope.PATCH
    wiener_weight_model: AggregationWeightModel = AggregationWeightModel.VARIANCE
    wiener_weight_domain: AggregationWeightDomain = AggregationWeightDomain.COEFFICIENT
    wiener_weight_scope: AggregationWeightScope = AggregationWeightScope.PATCH
    ht_use_soft_thresholding: bool = False
    ht_block_size: tuple[int, int] = (8, 8)
    ht_step: tuple[int, int] = (3, 3)
    ht_search_window: tuple[int, int] = (19, 19)
    ht_max_stack_size: int = 16
    ht_min_stack_size: int = 2
    ht_match_threshold: float = 2.9527
    ht_lambda_threshold: float = 3.0
    ht_kaiser_beta: float = 2.0
    wiener_block_size: tuple[int, int] = (8, 8)
    wiener_step: tuple[int, int] = (3, 3)
    wiener_search_window: tuple[int, int] = (19, 19)
    wiener_max_stack_size: int = 32
    wiener_min_stack_size: int = 2
    wiener_match_threshold: float = 0.3937
    wiener_variance_scale: float = 0.4
    wiener_kaiser_beta: float = 2.0
    ht_transform: Transform | list[Transform] = field(default_factory=_haar_bior)
    wiener_transform: Transform | list[Transform] = field(default_factory=_haar_dct)
    sharpen_alpha = 1.0
    sharpen_alpha_3d = 1.0
    ref_batch_size: int = 64


@dataclass
class BM4DProfile(BMNDProfile):
    """Official BM4D compatibility values (3-D volumes)."""

    noise_model: NoiseModel = NoiseModel.GAUSSIAN
    gamma: float = 3.0
    nf: tuple[int, int, int] | int | None = (16, 16, 16)
    k: int = 4
    ht_weight_model: AggregationWeightModel = AggregationWeightModel.VARIANCE
    ht_weight_domain: AggregationWeightDomain = AggregationWeightDomain.COEFFICIENT
    ht_weight_scope: AggregationWeightScope = AggregationWeightScope.PATCH
    wiener_weight_model: AggregationWeightModel = AggregationWeightModel.VARIANCE
    wiener_weight_domain: AggregationWeightDomain = AggregationWeightDomain.COEFFICIENT
    wiener_weight_scope: AggregationWeightScope = AggregationWeightScope.PATCH
    ht_use_soft_thresholding: bool = False
    ht_block_size: tuple[int, int, int] = (4, 4, 4)
    ht_step: tuple[int, int, int] = (3, 3, 3)
    ht_search_window: tuple[int, int, int] = (7, 7, 7)
    ht_max_stack_size: int = 16
    ht_min_stack_size: int = 2
    ht_match_threshold: float = 2.9527
    ht_lambda_threshold: float = 3.0
    ht_kaiser_beta: float = 2.0
    wiener_block_size: tuple[int, int, int] = (5, 5, 5)
    wiener_step: tuple[int, int, int] = (3, 3, 3)
    wiener_search_window: tuple[int, int, int] = (7, 7, 7)
    wiener_max_stack_size: int = 32
    wiener_min_stack_size: int = 2
    wiener_match_threshold: float = 0.7689
    wiener_variance_scale: float = 0.4
    wiener_kaiser_beta: float = 2.0
    ht_transform: Transform | list[Transform] = field(default_factory=_haar_bior)
    wiener_transform: Transform | list[Transform] = field(default_factory=_haar_dct)
    sharpen_alpha = 1.0
    sharpen_alpha_3d = 1.0
    ref_batch_size: int = 64


@dataclass
class BM5DProfile(BMNDProfile):
    """Four-dimensional BMND template."""

    noise_model: NoiseModel = NoiseModel.GAUSSIAN
    gamma: float = 3.0
    nf: tuple[int, int, int, int] | int | None = (16, 16, 16, 16)
    ht_use_soft_thresholding: bool = False
    ht_block_size: tuple[int, int, int, int] = (4, 4, 4, 4)
    ht_step: tuple[int, int, int, int] = (3, 3, 3, 3)
    ht_search_window: tuple[int, int, int, int] = (3, 3, 7, 7)
    ht_max_stack_size: int = 16
    ht_min_stack_size: int = 2
    ht_match_threshold: float = 2.9527
    ht_lambda_threshold: float = 3.0
    ht_kaiser_beta: float = 2.0
    wiener_block_size: tuple[int, int, int, int] = (4, 4, 4, 4)
    wiener_step: tuple[int, int, int, int] = (3, 3, 3, 3)
    wiener_search_window: tuple[int, int, int, int] = (3, 3, 7, 7)
    wiener_max_stack_size: int = 32
    wiener_min_stack_size: int = 2
    wiener_match_threshold: float = 0.7689
    wiener_variance_scale: float = 1.0
    wiener_kaiser_beta: float = 2.0
    ht_transform: Transform | list[Transform] = field(default_factory=_haar_bior)
    wiener_transform: Transform | list[Transform] = field(default_factory=_haar_dct)
    sharpen_alpha = 1.0
    sharpen_alpha_3d = 1.0
    ref_batch_size: int = 64
