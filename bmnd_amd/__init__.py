"""bmnd_amd -- the Poisson bmnd() denoiser on AMD Instinct MI355X (gfx950).

Drop-in for the reference package's public API (src/bmnd/__init__.py):
`bmnd`, the profile dataclasses and the string enums.  The work runs in
hand-written HIP kernels (libbmnd_hip.so) on one GPU.
"""

from .bmndalgo import bmnd, bmnd_batch
from .engine import set_exact_order
from .enums import (
    AggregationWeightDomain,
    AggregationWeightModel,
    AggregationWeightScope,
    BlockMatchDistance,
    CacheMode,
    NoiseModel,
    PoissonGroupMassConservation,
    PoissonHTMatchPolicy,
    PoissonVarianceSource,
    PoissonWienerGainMode,
    ReferenceScheduleMode,
)
from .profiles import BM2DProfile, BM3DProfile, BM4DProfile, BM5DProfile, BMNDProfile
from .transforms import Transform, TransformMode, TransformType

__all__ = [
    "AggregationWeightDomain",
    "AggregationWeightModel",
    "AggregationWeightScope",
    "BM2DProfile",
    "BM3DProfile",
    "BM4DProfile",
    "BM5DProfile",
    "BMNDProfile",
    "BlockMatchDistance",
    "CacheMode",
    "NoiseModel",
    "PoissonGroupMassConservation",
    "PoissonHTMatchPolicy",
    "PoissonVarianceSource",
    "PoissonWienerGainMode",
    "ReferenceScheduleMode",
    "Transform",
    "TransformMode",
    "TransformType",
    "bmnd",
    "bmnd_batch",
    "set_exact_order",
]
