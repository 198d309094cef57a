"""Patch geometry helpers (reference: src/bmnd/utils.py:5-31)."""

from __future__ import annotations

import numpy as np
from numpy.typing import NDArray


def kaiser_window_nd(patch_size: tuple[int, ...], beta: float = 2.0) -> NDArray[np.float32]:
    """Separable Kaiser window: outer product of np.kaiser in float64, cast
    to float32."""
    axes = [np.kaiser(int(sz), beta) for sz in patch_size]
    w = axes[0]
    for a in axes[1:]:
        w = np.multiply.outer(w, a)
    return np.asarray(w).astype(np.float32)


def extract_patches_strided(volume: NDArray[np.float32], patch_size: tuple[int, ...]):
    """All overlapping patches as a strided view (no copy)."""
    arr = np.asarray(volume)
    patch_size = tuple(patch_size)
    shape = tuple(arr.shape[d] - patch_size[d] + 1 for d in range(arr.ndim)) + patch_size
    return np.lib.stride_tricks.as_strided(arr, shape=shape, strides=arr.strides + arr.strides)
