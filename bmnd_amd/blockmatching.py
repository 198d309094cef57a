"""Block-matching result types and the threshold helpers of the bmnd API.

`BlockMatchGroup` / `BlockMatchResult` keep the reference's fields
(src/bmnd/blockmatching.py:151-165) and are built the same way as its
_build_blockmatch_group (blockmatching.py:1227-1262): absolute patch origins
from the linear indices and shifted positions relative to the group's
per-axis minimum.  They are only materialised for `blockmatch_callback`;
the device pipeline keeps match lists as flat int32 arrays.
"""

from __future__ import annotations

import functools
import math
from dataclasses import dataclass

import numpy as np
from numpy.typing import NDArray

from .enums import BlockMatchDistance, coerce_enum
from .shifts import ScheduleArrays
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


# ---- per-reference thresholds (poisson_ht_match_policy="reference_finite_count")

@functools.cache
def poisson_null_moment_table(distance_measure=BlockMatchDistance.POISSON_DEVIANCE,
                              max_count: int = 512) -> tuple[np.ndarray, np.ndarray]:
    """Mean and variance of one distance term when the reference and
    candidate counts are a Binomial(N, 1/2) split of a pooled count N, for N =
    0..max_count, summed in the reference's order in float64
    (blockmatching.py:29-87; N = 0 stays 0).  Pearson has the closed form
    mean 1, variance 2 - 2/N."""
    distance_measure = coerce_enum(distance_measure, BlockMatchDistance, "block-matching distance")
    if distance_measure not in (BlockMatchDistance.POISSON_DEVIANCE, BlockMatchDistance.PEARSON,
                                BlockMatchDistance.ANSCOMBE_SSD):
        raise ValueError("Poisson null moments require 'poisson_deviance', 'pearson', "
                         f"or 'anscombe_ssd', got {distance_measure!r}.")
    if max_count < 1:
        raise ValueError(f"max_count must be positive, got {max_count}.")
    means = np.zeros(max_count + 1, dtype=np.float64)
    variances = np.zeros(max_count + 1, dtype=np.float64)
    if distance_measure is BlockMatchDistance.PEARSON:
        n = np.arange(1, max_count + 1, dtype=np.float64)
        means[1:] = 1.0
        variances[1:] = 2.0 - 2.0 / n
        return means, variances
    deviance = distance_measure is BlockMatchDistance.POISSON_DEVIANCE
    for pooled in range(1, max_count + 1):
        expectation = 0.0
        second = 0.0
        scale = 2.0 ** (-pooled)
        for r in range(pooled + 1):
            c = pooled - r
            prob = math.comb(pooled, r) * scale
            if deviance:
                term = 0.0
                if r > 0:
                    term += r * math.log(2.0 * r / pooled)
                if c > 0:
                    term += c * math.log(2.0 * c / pooled)
                term *= 2.0
            else:  # Anscombe: squared difference of the stabilised counts / 2
                term = 0.5 * (2.0 * math.sqrt(c + 0.375) - 2.0 * math.sqrt(r + 0.375)) ** 2
            expectation += prob * term
            second += prob * term * term
        means[pooled] = expectation
        variances[pooled] = max(second - expectation * expectation, 0.0)
    return means, variances


def reference_patches(volume: np.ndarray, schedule: ScheduleArrays, block, start: int,
                      stop: int) -> np.ndarray:
    """float32 patches (rows in C order) of references start..stop-1."""
    ref = schedule.reference_abs[start:stop]
    offs = np.stack(np.meshgrid(*[np.arange(b) for b in block], indexing="ij"),
                    axis=-1).reshape(-1, len(block))
    idx = ref[:, None, :] + offs[None, :, :]
    return volume[tuple(idx[..., d] for d in range(len(block)))].astype(np.float32, copy=False)


def compute_poisson_reference_thresholds(volume: np.ndarray, schedule: ScheduleArrays, block,
                                         count_scale: float, ht_match_threshold: float,
                                         structure_beta: float, intensity_power: float,
                                         max_count: int = 512, chunk: int = 65536,
                                         distance=BlockMatchDistance.POISSON_DEVIANCE) -> np.ndarray:
    """Per-reference stage-1 match thresholds of the "reference_finite_count"
    policy (blockmatching.py:95-148), vectorised over references: the mean
    of the reference patch's null means plus ht_match_threshold times the
    standard deviation of the null distance, plus structure_beta times the
    mean reference count to intensity_power; float32 per reference."""
    count_scale = float(count_scale)
    if not np.isfinite(ht_match_threshold) or ht_match_threshold < 0.0:
        raise ValueError("ht_match_threshold must be finite and non-negative, "
                         f"got {ht_match_threshold}.")
    if not np.isfinite(structure_beta) or structure_beta < 0.0:
        raise ValueError(f"structure_beta must be finite and non-negative, got {structure_beta}.")
    if not np.isfinite(intensity_power) or intensity_power <= 0.0:
        raise ValueError(f"intensity_power must be finite and positive, got {intensity_power}.")
    means, variances = poisson_null_moment_table(distance, max_count)
    P = int(np.prod(block))
    out = np.empty(len(schedule), dtype=np.float32)
    for a in range(0, len(schedule), chunk):
        b = min(a + chunk, len(schedule))
        counts = np.clip(reference_patches(volume, schedule, block, a, b).astype(np.float64)
                         / count_scale, 0.0, None)
        pooled = np.rint(2.0 * counts)
        idx = np.minimum(pooled, max_count).astype(np.int64)
        null_means = means[idx]
        null_vars = variances[idx]
        above = pooled > max_count
        null_means[above] = 1.0
        null_vars[above] = 2.0
        patch_mean = np.mean(null_means, axis=1)
        patch_var = np.sum(null_vars, axis=1) / (P * P)
        intensity = np.mean(counts, axis=1)
        thr = (patch_mean + float(ht_match_threshold) * np.sqrt(np.maximum(patch_var, 0.0))
               + float(structure_beta) * np.power(intensity, float(intensity_power)))
        out[a:b] = thr.astype(np.float32)
    return out


def poisson_null_moment_table_f32(distance_measure, max_count: int = 512) -> np.ndarray:
    """The null table as the candidate-standardized distance reads it: the
    float32 means for pooled counts 0..max_count, then the float32 variances
    (blockmatching.py:288-319)."""
    if isinstance(max_count, bool) or not isinstance(max_count, (int, np.integer)) \
            or int(max_count) <= 0:
        raise ValueError(
            f"poisson_moment_max_count must be a positive integer, got {max_count!r}.")
    m, v = poisson_null_moment_table(distance_measure, int(max_count))
    return np.concatenate([m.astype(np.float32), v.astype(np.float32)])
