"""Block-matching result types and the threshold helpers of the bmnd API.

`BlockMatchGroup` / `BlockMatchResult` keep the reference's fields
(src/bmnd/blockmatching.py:151-165) and are built the same way as its
_build_blockmatch_group (blockmatching.py:1227-1262): absolute patch origins
from the linear indices and shifted positions relative to the group's
per-axis minimum.  They are only materialised for `blockmatch_callback`;
the device pipeline keeps match lists as flat int32 arrays.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from numpy.typing import NDArray

from .enums import BlockMatchDistance, coerce_enum
from .profilevalidation import resolve_blockmatch_distance  # noqa: F401  (re-export)


@dataclass
class BlockMatchGroup:
    reference_abs: tuple[int, ...]
    reference_shift: tuple[int, ...]
    reference_shifted: tuple[int, ...]
    selected_lin_indices: NDArray[np.int64]
    selected_abs_positions: NDArray[np.int64]
    selected_shifted_positions: NDArray[np.int64]
    accepted_count: int
    final_count: int


@dataclass
class BlockMatchResult:
    groups: list[BlockMatchGroup]


def compute_blockmatch_threshold(similarity_threshold: float, volume,
                                 distance_measure=BlockMatchDistance.SSD) -> float:
    """Deviance thresholds are used as given; SSD thresholds scale with the
    squared data range (blockmatching.py:168-184)."""
    distance_measure = coerce_enum(distance_measure, BlockMatchDistance, "block-matching distance")
    if distance_measure is not BlockMatchDistance.SSD:
        return float(similarity_threshold)
    arr = np.asarray(volume, dtype=np.float32)
    if arr.size == 0:
        return float(similarity_threshold)
    scale = float(np.max(arr) - np.min(arr))
    if scale <= 0.0:
        scale = 1.0
    return float(similarity_threshold) * scale * scale


def build_result(schedule, counts: tuple[int, ...], sel: np.ndarray, final: np.ndarray,
                 accepted: np.ndarray, include_positions: bool = True) -> BlockMatchResult:
    """Assemble BlockMatchGroup objects from flat device outputs."""
    ndim = len(counts)
    groups: list[BlockMatchGroup] = []
    for g in range(len(final)):
        lins = sel[g, : int(final[g])].astype(np.int64)
        if include_positions and lins.size:
            absp = np.stack(np.unravel_index(lins, counts), axis=1).astype(np.int64)
            shifted = absp - absp.min(axis=0, keepdims=True)
            ref_shifted = tuple(int(x) for x in shifted[0])
        else:
            absp = np.empty((0, ndim), dtype=np.int64)
            shifted = absp.copy()
            ref_shifted = tuple(int(x) for x in schedule.reference_shifted[g])
        groups.append(BlockMatchGroup(
            reference_abs=tuple(int(x) for x in schedule.reference_abs[g]),
            reference_shift=tuple(int(x) for x in schedule.shift[g]),
            reference_shifted=ref_shifted,
            selected_lin_indices=lins,
            selected_abs_positions=absp,
            selected_shifted_positions=shifted,
            accepted_count=int(accepted[g]),
            final_count=int(final[g])))
    return BlockMatchResult(groups=groups)
