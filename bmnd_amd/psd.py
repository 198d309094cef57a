"""Stationary-noise description of the Gaussian path, on the host: the noise
PSD from a scalar sigma or an explicit sigma_psd, its preprocessing for
the nf variance domain, the autocovariance, and the Gaussian global sigma
estimate.  Restates src/bmnd/psd.py:19-261 and sigmaestimation.py:7-45 in
the same numpy / scipy steps, so the tables the device kernels read carry
the reference's values.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.ndimage import correlate


def validate_sigma(sigma) -> float:
    """psd.py:19-30."""
    if isinstance(sigma, (bool, np.bool_)):
        raise ValueError(f"sigma must be finite and non-negative, got {sigma!r}.")
    try:
        value = float(sigma)
    except (TypeError, ValueError) as exc:
        raise ValueError(f"sigma must be finite and non-negative, got {sigma!r}.") from exc
    if not np.isfinite(value) or value < 0.0:
        raise ValueError(f"sigma must be finite and non-negative, got {sigma!r}.")
    return value


def validate_sigma_psd(shape, sigma_psd) -> np.ndarray:
    """psd.py:33-55: finite, non-negative, scalar or of the volume's shape."""
    try:
        with np.errstate(over="ignore", invalid="ignore"):
            arr = np.asarray(sigma_psd, dtype=np.float32)
    except (TypeError, ValueError) as exc:
        raise ValueError("sigma_psd must contain only finite non-negative values.") from exc
    if not np.all(np.isfinite(arr)) or np.any(arr < 0.0):
        raise ValueError("sigma_psd must contain only finite non-negative values.")
    if arr.ndim != 0 and arr.size != 1 and arr.shape != tuple(shape):
        raise ValueError("sigma_psd must either be scalar or have the same shape as volume")
    return arr


def resolve_nf(nf, shape):
    """psd.py:58-99: (the nf domain or None, whether nf asked for the
    conventional path)."""
    if nf is None:
        return None, False
    if isinstance(nf, bool):
        raise ValueError(f"nf must be a per-axis integer tuple or zero, got {nf!r}.")
    if isinstance(nf, (int, np.integer)):
        if int(nf) == 0:
            return tuple(min(int(s), 16) for s in shape), True
        raise ValueError("nf must be zero or contain one positive integer per volume axis, "
                         f"got {nf!r}.")
    try:
        values = tuple(nf)
    except TypeError as exc:
        raise ValueError("nf must be zero or contain one positive integer per volume axis, "
                         f"got {nf!r}.") from exc
    if not values:
        return None, False
    if len(values) != len(shape):
        raise ValueError(f"nf must contain {len(shape)} values for shape {shape}, got {values}.")
    if any(isinstance(v, bool) or not isinstance(v, (int, np.integer)) for v in values):
        raise ValueError(f"nf values must be integers, got {values}.")
    resolved = tuple(int(v) for v in values)
    if any(v < 0 for v in resolved):
        raise ValueError(f"nf values must be non-negative, got {resolved}.")
    if any(v == 0 for v in resolved):
        return tuple(min(int(s), 16) for s in shape), True
    return resolved, False


def scalar_sigma_to_psd(shape, sigma: float) -> np.ndarray:
    """psd.py:102-112: the flat PSD N sigma^2 (N the volume size)."""
    value = float(np.prod(shape)) * validate_sigma(sigma) ** 2
    if not np.isfinite(value) or value > np.finfo(np.float32).max:
        raise ValueError(f"sigma is too large for a float32 PSD, got {sigma!r}.")
    return np.full(tuple(shape), value, dtype=np.float32)


def normalize_sigma_psd(shape, sigma=None, sigma_psd=None):
    """psd.py:115-131."""
    if sigma is not None and sigma_psd is not None:
        raise ValueError("sigma and sigma_psd are mutually exclusive.")
    if sigma is None and sigma_psd is None:
        return None
    if sigma_psd is None:
        return scalar_sigma_to_psd(shape, sigma)
    arr = validate_sigma_psd(shape, sigma_psd)
    if arr.ndim == 0 or arr.size == 1:
        return scalar_sigma_to_psd(shape, float(np.asarray(arr).reshape(())))
    return arr.astype(np.float32, copy=False)


def autocovariance_from_psd(sigma_psd) -> np.ndarray:
    """psd.py:134-139: real(ifftn(PSD)) / N in float32."""
    psd = np.asarray(sigma_psd, dtype=np.float32)
    return np.real(np.fft.ifftn(psd) / float(psd.size)).astype(np.float32, copy=False)


def kernel_from_psd(sigma_psd, single_dim_psd: bool = False) -> np.ndarray:
    """psd.py:142-156."""
    psd = np.asarray(sigma_psd, dtype=np.float32)
    if single_dim_psd:
        return np.array([float(np.linalg.norm(np.sqrt(psd / float(psd.size))))], np.float32)
    std = np.sqrt(psd / float(psd.size))
    return np.fft.fftshift(np.real(np.fft.ifftn(std)).astype(np.float32)).astype(np.float32)


def gaussian_kernel_nd(size, sigma_scale) -> np.ndarray:
    """psd.py:159-176: normalised separable Gaussian, float32."""
    axes = []
    for n, s in zip(size, sigma_scale):
        coords = np.arange(n, dtype=np.float32) - (n - 1) / 2.0
        axis = np.exp(-0.5 * (coords / max(float(s), 1e-6)) ** 2)
        axis /= np.sum(axis, dtype=np.float32)
        axes.append(axis.astype(np.float32))
    kernel = axes[0]
    for axis in axes[1:]:
        kernel = np.multiply.outer(kernel, axis)
    kernel = kernel.astype(np.float32, copy=False)
    kernel /= np.sum(kernel, dtype=np.float32)
    return kernel


def nf_blur_kernel(psd_shape, nf):
    """psd.py:189-201."""
    sizes, sigmas = [], []
    for axis in range(len(psd_shape)):
        half = np.floor(0.5 * psd_shape[axis] / nf[axis])
        sizes.append(int(1 + 2 * half))
        sigmas.append(float(1 + 2 * half / 20.0))
    return gaussian_kernel_nd(tuple(sizes), tuple(sigmas)), tuple(sigmas)


def reduce_and_blur_psd(sigma_psd, nf):
    """psd.py:204-241: (reduced PSD, blur kernel, blurred PSD); the PSD is
    decimated by 3-tap means (wrap) until each axis is within 16x of nf,
    rescaled, then blurred with the nf kernel (wrap)."""
    psd = np.array(sigma_psd, dtype=np.float32, copy=True)
    if nf is None:
        return psd, np.ones((1,) * psd.ndim, np.float32), psd.copy()
    reduced = psd.copy()
    taps = []
    for axis in range(psd.ndim):
        shape = [1] * psd.ndim
        shape[axis] = 3
        taps.append(np.ones(tuple(shape), np.float32) / np.float32(3.0))
    for axis in range(psd.ndim):
        original = psd.shape[axis] / float(nf[axis])
        ratio = reduced.shape[axis] / float(nf[axis])
        while ratio > 16.0:
            mid = correlate(reduced, taps[axis], mode="wrap")
            sl = [slice(None)] * psd.ndim
            sl[axis] = slice(1, None, 3)
            reduced = mid[tuple(sl)]
            ratio = reduced.shape[axis] / float(nf[axis])
        reduced *= np.float32(ratio / original)
    blur, _ = nf_blur_kernel(tuple(int(s) for s in reduced.shape), nf)
    blurred = correlate(reduced, blur, mode="wrap").astype(np.float32, copy=False)
    return reduced.astype(np.float32, copy=False), blur, blurred


@dataclass
class ProcessedPSD:
    sigma_psd: np.ndarray
    correlation_kernel: np.ndarray
    reduced_sigma_psd: np.ndarray
    blur_kernel: np.ndarray
    blurred_sigma_psd: np.ndarray


def preprocess_psd(sigma_psd, nf, single_dim_psd: bool = False) -> ProcessedPSD:
    """psd.py:244-261."""
    psd = np.array(sigma_psd, dtype=np.float32, copy=True)
    resolved, _ = resolve_nf(nf, tuple(int(s) for s in psd.shape))
    kernel = kernel_from_psd(psd, single_dim_psd=single_dim_psd)
    reduced, blur, blurred = reduce_and_blur_psd(psd, None if single_dim_psd else resolved)
    return ProcessedPSD(psd, kernel, reduced, blur, blurred)


def estimate_gaussian_global_sigma(volume, eps: float = 1e-10) -> float:
    """sigmaestimation.py:7-45: MAD / 0.6745 of the finite non-zero values,
    then the IQR, the trimmed std and sqrt(eps) as fallbacks."""
    vol = np.asarray(volume, dtype=np.float32)
    finite = np.isfinite(vol)
    nonzero = finite & (vol != 0.0)
    if np.any(nonzero):
        ref = vol[nonzero]
    elif np.any(finite):
        ref = vol[finite]
    else:
        ref = np.array([0.0], dtype=np.float32)
    med = np.median(ref)
    sigma = float(np.median(np.abs(ref - med))) / 0.6745
    if sigma <= 0.0:
        q25, q75 = np.percentile(ref, [25.0, 75.0])
        iqr = float(q75 - q25)
        if iqr > 0.0:
            sigma = iqr / 1.349
    if sigma <= 0.0:
        p5, p95 = np.percentile(ref, [5.0, 95.0])
        trimmed = ref[(ref >= p5) & (ref <= p95)]
        if trimmed.size > 0:
            s = float(np.std(trimmed, dtype=np.float32))
            if s > 0.0:
                sigma = s
    if sigma <= 0.0:
        sigma = float(np.sqrt(eps))
    if not np.isfinite(sigma) or sigma < 0.0:
        return 0.0
    return sigma
