"""Stationary Gaussian variance model of the Gaussian bmnd(), on the host:
the model (global sigma, processed PSD, autocovariance, nf domain), the
block-matching noise bias per search lag and the per-coefficient
cross-covariance maps over the nf-periodic domain that the device's group
variance kernel reads.  These are small per-(profile, sigma) tables, built
with the reference's numpy / scipy steps (variance.py:452-553, 1714-1886) so
that their values are the reference's; the per-group work that consumes
them runs on the device.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy.ndimage import map_coordinates

from .psd import (
    autocovariance_from_psd,
    estimate_gaussian_global_sigma,
    normalize_sigma_psd,
    preprocess_psd,
    resolve_nf,
)


@dataclass
class GaussianVarianceModel:
    global_sigma: float
    sigma_psd: np.ndarray
    reduced_sigma_psd: np.ndarray
    processed_sigma_psd: np.ndarray
    blur_kernel: np.ndarray
    correlation_kernel: np.ndarray
    autocovariance: np.ndarray
    variance_domain_shape: tuple[int, ...]
    _maps: dict = field(default_factory=dict)
    _ssd: dict = field(default_factory=dict)


def build_variance_model(volume, sigma=None, sigma_psd=None, nf=None) -> GaussianVarianceModel:
    """variance.py:452-504: sigma estimated from the volume (MAD) when
    neither sigma nor sigma_psd is given; a scalar input skips the PSD
    reduction; the variance domain is the resolved nf, else the PSD's shape."""
    if sigma is not None and sigma_psd is not None:
        raise ValueError("sigma and sigma_psd are mutually exclusive.")
    if sigma is None and sigma_psd is None:
        sigma = estimate_gaussian_global_sigma(volume)
    shape = tuple(int(d) for d in np.shape(volume))
    resolved, _ = resolve_nf(nf, shape)
    psd = normalize_sigma_psd(shape, sigma=sigma, sigma_psd=sigma_psd)
    if sigma is not None:
        scalar = True
    else:
        arr = np.asarray(sigma_psd)
        scalar = arr.ndim == 0 or arr.size == 1
    proc = preprocess_psd(psd, resolved, single_dim_psd=scalar)
    processed = proc.blurred_sigma_psd
    mean = float(np.mean(psd, dtype=np.float32))
    global_sigma = float(np.sqrt(max(mean / float(np.prod(shape)), 0.0)))
    domain = tuple(int(s) for s in processed.shape) if resolved is None else tuple(resolved)
    return GaussianVarianceModel(
        global_sigma=max(global_sigma, 0.0), sigma_psd=psd,
        reduced_sigma_psd=proc.reduced_sigma_psd, processed_sigma_psd=processed,
        blur_kernel=proc.blur_kernel, correlation_kernel=proc.correlation_kernel,
        autocovariance=autocovariance_from_psd(processed), variance_domain_shape=domain)


def blockmatch_noise_ssd(model: GaussianVarianceModel, block, radius) -> np.ndarray:
    """variance.py:507-553: the expected SSD of two noise patches at each
    search lag, 2 P (var - lagcov) with lagcov the symmetrised
    autocovariance, 0 at the zero lag; float32 of shape 2 r + 1."""
    ac = model.autocovariance
    nd = ac.ndim
    block, radius = tuple(int(b) for b in block), tuple(int(r) for r in radius)
    if len(block) != nd or any(b <= 0 for b in block):
        raise ValueError(f"block_shape must contain {nd} positive values, got {block}.")
    if len(radius) != nd or any(r < 0 for r in radius):
        raise ValueError(f"search_radius must contain {nd} non-negative values, got {radius}.")
    key = (block, radius)
    if key in model._ssd:
        return model._ssd[key]
    var = float(ac[(0,) * nd])
    pvol = float(np.prod(block))
    # every lag at once: the same float64 expression per lag, then float32
    lags = np.meshgrid(*[np.arange(-r, r + 1) for r in radius], indexing="ij")
    pos = tuple(lag % ac.shape[a] for a, lag in enumerate(lags))
    neg = tuple((-lag) % ac.shape[a] for a, lag in enumerate(lags))
    lagcov = 0.5 * (ac[pos] + ac[neg]).astype(np.float64)
    out = np.maximum(2.0 * pvol * (var - lagcov), 0.0).astype(np.float32)
    out[radius] = np.float32(0.0)
    model._ssd[key] = out
    return out


def _kron_chain(block, spatial_forward) -> np.ndarray:
    """variance.py:744-756: the float32 Kronecker chain of the spatial
    analysis matrices."""
    full = np.array([[1.0]], dtype=np.float32)
    for b, m in zip(block, spatial_forward):
        mat = np.eye(b, dtype=np.float32) if m is None else np.asarray(m, dtype=np.float32)
        full = np.kron(full, mat).astype(np.float32, copy=False)
    return full


def _domain_psd(model: GaussianVarianceModel) -> np.ndarray:
    """The processed PSD on the variance domain: resampled (linear, periodic)
    and rescaled when its shape differs (variance.py:1823-1843)."""
    psd = np.asarray(model.processed_sigma_psd, dtype=np.float32)
    domain = model.variance_domain_shape
    if psd.shape == domain:
        return psd
    src = tuple(int(s) for s in psd.shape)
    axes = [np.arange(t, dtype=np.float64) * (s / float(t)) for s, t in zip(src, domain)]
    coords = np.stack(np.meshgrid(*axes, indexing="ij"), axis=0)
    out = map_coordinates(psd, coords, order=1, mode="grid-wrap").astype(np.float32, copy=False)
    out *= np.float32(np.prod(domain) / float(np.prod(src)))
    return np.maximum(out, 0.0)


def cross_covariance_maps(model: GaussianVarianceModel, block, spatial_forward) -> np.ndarray:
    """variance.py:1791-1886: for every spatial coefficient s, the noise
    covariance of coefficient s between two patches as a function of their
    lag on the nf-periodic domain: the analysis row of s folded onto the
    domain, its power spectrum times the PSD, inverse-transformed and
    symmetrised in the lag.  Float32, shape (P,) + domain.  Entry 0 of each
    map is the coefficient's variance."""
    block = tuple(int(b) for b in block)
    domain = tuple(int(d) for d in model.variance_domain_shape)
    if len(domain) != len(block):
        raise ValueError("Processed PSD rank does not match block rank: "
                         f"got {domain} and {block}.")
    key = (block, tuple(np.asarray(m, np.float32).tobytes() for m in spatial_forward))
    if key in model._maps:
        return model._maps[key]
    spatial = _kron_chain(block, spatial_forward)
    psd = _domain_psd(model)
    P = int(np.prod(block))
    local = np.stack(np.meshgrid(*[np.arange(b, dtype=np.int64) for b in block], indexing="ij"),
                     axis=-1).reshape(-1, len(block))
    folded_index = np.ravel_multi_index(tuple((local % np.array(domain)).T), domain)
    rev = np.ix_(*[(-np.arange(d, dtype=np.int64)) % d for d in domain])
    maps = np.empty((P,) + domain, dtype=np.float32)
    for s in range(P):
        folded = np.zeros(int(np.prod(domain)), dtype=np.float32)
        # float32 adds in local-index order (several entries share a cell
        # only when the block is longer than the domain on an axis)
        np.add.at(folded, folded_index, spatial[s])
        spec = np.fft.fftn(folded.reshape(domain))
        cmap = np.real(np.fft.ifftn(psd * (np.abs(spec) ** 2)) / float(psd.size)).astype(
            np.float32, copy=False)
        maps[s] = 0.5 * (cmap + cmap[rev])
    model._maps[key] = maps
    return maps

