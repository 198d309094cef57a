"""String enums of the bmnd API (same names and values as the reference's
src/bmnd/enums.py:5-70, so profiles written for it load unchanged)."""

from __future__ import annotations

from enum import Enum
from typing import TypeVar


class StrEnum(str, Enum):
    """`enum.StrEnum` for Python 3.10: members compare and print as their value."""

    def __str__(self) -> str:
        return str(self.value)

    def __format__(self, spec: str) -> str:
        return format(str(self.value), spec)


class NoiseModel(StrEnum):
    GAUSSIAN = "gaussian"
    POISSON = "poisson"


class ReferenceScheduleMode(StrEnum):
    OFFICIAL = "official"
    GENERATED = "generated"
    BALANCED = "balanced"
    SPARSE = "sparse"
    OFF = "off"


class CacheMode(StrEnum):
    AUTO = "auto"


class BlockMatchDistance(StrEnum):
    AUTO = "auto"
    SSD = "ssd"
    POISSON_DEVIANCE = "poisson_deviance"
    PEARSON = "pearson"
    ANSCOMBE_SSD = "anscombe_ssd"


class AggregationWeightModel(StrEnum):
    CLASSIC = "classic"
    VARIANCE = "variance"
    RISK = "risk"


class AggregationWeightDomain(StrEnum):
    COEFFICIENT = "coefficient"
    WINDOWED_SYNTHESIS = "windowed_synthesis"


class AggregationWeightScope(StrEnum):
    GROUP = "group"
    PATCH = "patch"


class PoissonVarianceSource(StrEnum):
    OBSERVATION = "observation"
    PILOT = "pilot"


class PoissonHTMatchPolicy(StrEnum):
    FIXED = "fixed"
    REFERENCE_FINITE_COUNT = "reference_finite_count"
    CANDIDATE_STANDARDIZED = "candidate_standardized"


class PoissonWienerGainMode(StrEnum):
    CLASSIC = "classic"
    NOISE_FLOOR = "noise-floor"
    VARIANCE_SCALED = "variance-scaled"


class PoissonGroupMassConservation(StrEnum):
    NONE = "none"
    HT = "ht"
    WIENER = "wiener"
    BOTH = "both"


EnumT = TypeVar("EnumT", bound=Enum)


def coerce_enum(value: EnumT | str, enum_type: type[EnumT], name: str) -> EnumT:
    """Accept an enum member or its string value (enums.py:73-80 semantics)."""
    if isinstance(value, enum_type):
        return value
    try:
        return enum_type(value)
    except (TypeError, ValueError) as exc:
        expected = ", ".join(repr(member.value) for member in enum_type)
        raise ValueError(f"Unknown {name}: {value!r} (expected {expected}).") from exc
