"""Inputs and oracle set-up of the stationary Gaussian tests: seeded noisy
phantoms, and the oracle's parameters for a profile with the noise
description (host tables) of the product's Gaussian pipeline."""

from __future__ import annotations

import numpy as np

import oracle
from bmnd_amd import BM3DProfile, BM4DProfile, phantoms
from bmnd_amd.gaussianvariance import blockmatch_noise_ssd, cross_covariance_maps
from bmnd_amd.profilevalidation import resolve_profile
from bmnd_amd.transforms import spatial_matrices
from bmnd_amd.engine import make_plan


def noisy_input(kind: str, sigma: float = 0.1, seed: int = 0, shape=None):
    """(truth, noisy, profile): a smooth phantom in [0, 1] plus white
    Gaussian noise of std `sigma`."""
    rng = np.random.default_rng(seed)
    if kind == "volume":
        shp = shape or (24, 26, 22)
        truth, _ = phantoms.ellipsoid_volume(shp, seed=seed + 5, peak=1.0, n_ellipsoids=6,
                                             axes_range=(3.0, 10.0))
        prof = BM4DProfile()
    else:
        size = shape[0] if shape else 64
        truth, _ = phantoms.blob_image(size, seed=seed + 5, peak=1.0)
        prof = BM3DProfile()
    truth = np.asarray(truth, np.float32)
    truth = truth / max(float(truth.max()), 1e-6)
    noisy = (truth + sigma * rng.standard_normal(truth.shape)).astype(np.float32)
    return truth, noisy, prof


def oracle_params(profile, shape, model, nthreads: int = 0):
    """The oracle's Gaussian parameters for `profile` on `shape` with the
    noise model `model` (gaussianvariance.build_variance_model, or the
    `model` of GaussianPipeline.tables())."""
    shape = tuple(int(s) for s in shape)
    ndim = len(shape)
    resolved = resolve_profile(profile, shape)
    p = oracle.default_params(ndim, shape, nthreads)
    maps = {}
    for stage in ("ht", "wiener"):
        plan = make_plan(profile, stage, ndim)
        mats = [spatial_matrices(plan.spatial, b)[0] for b in plan.block]
        maps[stage] = cross_covariance_maps(model, plan.block, mats)
    gamma = 0.0 if resolved.conventional_nf else float(profile.gamma)
    noise = None
    if gamma != 0.0:
        ht = make_plan(profile, "ht", ndim)
        noise = blockmatch_noise_ssd(model, ht.block, ht.radius3[3 - ndim:])
    oracle.set_gaussian(p, global_sigma=model.global_sigma,
                        domain=model.variance_domain_shape, ht_maps=maps["ht"],
                        wi_maps=maps["wiener"], ht_match=float(profile.ht_match_threshold),
                        wi_match=float(profile.wiener_match_threshold), ht_noise_ssd=noise,
                        gamma=gamma, k=resolved.variance_k)
    p.ht_lambda = float(profile.ht_lambda_threshold)
    p.wi_mu = float(profile.wiener_variance_scale)
    return p
