"""Synthetic Poisson workloads (seeded).

Generators for the benchmark and test inputs: the README square image, the
2-D Gaussian-blob-and-rectangle images and the 3-D ellipsoid volumes, each
scaled to a peak count and Poisson-sampled with numpy's default_rng(seed).
"""

from __future__ import annotations

import numpy as np


def square_image(seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """64x64 intensity 5 with a 32x32 square of 20 at [16:48, 16:48]."""
    img = np.full((64, 64), 5.0, dtype=np.float32)
    img[16:48, 16:48] = 20.0
    noisy = np.random.default_rng(seed).poisson(img).astype(np.float32)
    return img, noisy


def blob_image(size: int = 512, seed: int = 0, peak: float | None = None,
               n_gauss: int = 20, n_rect: int = 10) -> tuple[np.ndarray, np.ndarray]:
    """Sum of anisotropic Gaussians (sigma 4-60 px) and axis-aligned
    rectangles, rescaled to `peak` counts (drawn from {1,5,20,100} when None)."""
    rng = np.random.default_rng(seed)
    if peak is None:
        peak = float(rng.choice([1.0, 5.0, 20.0, 100.0]))
    yy, xx = np.mgrid[0:size, 0:size].astype(np.float64)
    img = np.zeros((size, size), dtype=np.float64)
    for _ in range(n_gauss):
        cy, cx = rng.uniform(0, size, 2)
        sy, sx = rng.uniform(4, 60, 2)
        amp = rng.uniform(0.2, 1.0)
        img += amp * np.exp(-0.5 * (((yy - cy) / sy) ** 2 + ((xx - cx) / sx) ** 2))
    for _ in range(n_rect):
        y0, x0 = rng.integers(0, size - 8, 2)
        h, w = rng.integers(8, max(9, size // 4), 2)
        img[y0:y0 + h, x0:x0 + w] += rng.uniform(0.2, 1.0)
    img = img / img.max() * peak
    noisy = rng.poisson(img).astype(np.float32)
    return img.astype(np.float32), noisy


def ellipsoid_volume(shape=(256, 256, 256), seed: int = 0, peak: float = 10.0,
                     n_ellipsoids: int = 40, floor: float = 0.05,
                     axes_range=(5.0, 60.0)) -> tuple[np.ndarray, np.ndarray]:
    """Additive ellipsoids (semi-axes in `axes_range`, intensity 0.2-1.0)
    plus a floor, scaled to `peak` counts and Poisson-sampled."""
    rng = np.random.default_rng(seed)
    shape = tuple(int(s) for s in shape)
    vol = np.full(shape, floor, dtype=np.float32)
    grids = [np.arange(s, dtype=np.float32) for s in shape]
    for _ in range(n_ellipsoids):
        c = [rng.uniform(0, s) for s in shape]
        ax = rng.uniform(axes_range[0], axes_range[1], 3)
        inten = np.float32(rng.uniform(0.2, 1.0))
        lo = [max(int(c[d] - ax[d]), 0) for d in range(3)]
        hi = [min(int(c[d] + ax[d]) + 1, shape[d]) for d in range(3)]
        if any(hi[d] <= lo[d] for d in range(3)):
            continue
        z = ((grids[0][lo[0]:hi[0]] - c[0]) / ax[0]) ** 2
        y = ((grids[1][lo[1]:hi[1]] - c[1]) / ax[1]) ** 2
        x = ((grids[2][lo[2]:hi[2]] - c[2]) / ax[2]) ** 2
        inside = (z[:, None, None] + y[None, :, None] + x[None, None, :]) <= 1.0
        vol[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] += inten * inside
    vol = vol / vol.max() * np.float32(peak)
    noisy = rng.poisson(vol).astype(np.float32)
    return vol.astype(np.float32), noisy


def psnr(estimate: np.ndarray, truth: np.ndarray, peak: float | None = None) -> float:
    peak = float(np.max(truth)) if peak is None else float(peak)
    mse = float(np.mean((np.asarray(estimate, np.float64) - np.asarray(truth, np.float64)) ** 2))
    return float("inf") if mse == 0 else 10.0 * np.log10(peak * peak / mse)


def uneven_batch(seed: int = 0, n_images: int = 32, n_volumes: int = 16,
                 image_sides=(96, 1024), volume_sides=(48, 128), peaks=(1.0, 50.0)):
    """Inputs of uneven shapes: images with each side uniform in `image_sides`
    (blob phantoms cropped from a square one) and volumes with each side
    uniform in `volume_sides` (ellipsoid phantoms), every input at a peak
    count uniform in `peaks`.  Returns (images, image_truths, volumes,
    volume_truths)."""
    rng = np.random.default_rng(seed)
    imgs, itruth, vols, vtruth = [], [], [], []
    for _ in range(n_images):
        h, w = (int(x) for x in rng.integers(image_sides[0], image_sides[1] + 1, 2))
        truth, noisy = blob_image(max(h, w), seed=int(rng.integers(1 << 30)),
                                  peak=float(rng.uniform(*peaks)))
        itruth.append(np.ascontiguousarray(truth[:h, :w]))
        imgs.append(np.ascontiguousarray(noisy[:h, :w]))
    for _ in range(n_volumes):
        shape = tuple(int(x) for x in rng.integers(volume_sides[0], volume_sides[1] + 1, 3))
        truth, noisy = ellipsoid_volume(shape, seed=int(rng.integers(1 << 30)),
                                        peak=float(rng.uniform(*peaks)))
        vtruth.append(truth)
        vols.append(noisy)
    return imgs, itruth, vols, vtruth
