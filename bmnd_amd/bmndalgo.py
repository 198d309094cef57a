"""bmnd(): the drop-in entry point.

Same signature, argument validation order, error messages and return
conventions as the reference's src/bmnd/bmndalgo.py:71-183; the work runs on
the MI355X device pipelines (engine.PoissonPipeline for Poisson noise,
gaussian.GaussianPipeline for stationary Gaussian noise).  Profile options
the device pipelines do not run (engine.check_supported: the exact
covariance-aware group risk under Gaussian noise, non-SSD distances under
Gaussian noise, transforms and geometries not compiled in) raise
NotImplementedError after the reference's own validation has passed.
"""

from __future__ import annotations

from collections.abc import Callable

import numpy as np
import torch

from .blockmatching import BlockMatchResult, build_result
from .enums import CacheMode, coerce_enum
from .engine import BatchPipeline, PoissonPipeline, check_supported, table_first_order
from .gaussian import GaussianPipeline
from .profilevalidation import resolve_profile
from .psd import validate_sigma_psd


def _validate_sigma(sigma: object) -> float:
    msg = f"sigma must be finite and non-negative, got {sigma!r}."
    if isinstance(sigma, (bool, np.bool_)):
        raise ValueError(msg)
    try:
        value = float(sigma)
    except (TypeError, ValueError) as exc:
        raise ValueError(msg) from exc
    if not np.isfinite(value) or value < 0.0:
        raise ValueError(msg)
    return value


def _validate_max_cache_bytes(max_cache_bytes):
    """Same contract as the reference (cache.py:65-83).  The device pipeline
    keeps no covariance caches, so the value has no further effect."""
    if isinstance(max_cache_bytes, str):
        max_cache_bytes = coerce_enum(max_cache_bytes, CacheMode, "max_cache_bytes mode")
    if max_cache_bytes is None or max_cache_bytes is CacheMode.AUTO:
        return max_cache_bytes
    if (isinstance(max_cache_bytes, bool) or not isinstance(max_cache_bytes, (int, np.integer))
            or int(max_cache_bytes) < 0):
        raise ValueError("max_cache_bytes must be a non-negative integer, 'auto', or None, "
                         f"got {max_cache_bytes!r}.")
    return int(max_cache_bytes)


_PIPELINES: dict = {}

# bmnd() runs both stages as one captured HIP graph (engine.StageGraph) for
# inputs up to this many voxels, where launches and read-backs dominate, when
# no callback or sigma map is asked for; the result is the same bit for bit.
USE_GRAPHS = True
GRAPH_MAX_VOXELS = 1 << 22


def _device() -> torch.device:
    """The caller's current HIP device."""
    return torch.device("cuda", torch.cuda.current_device())


def _pipeline(profile, shape, device) -> PoissonPipeline:
    key = (repr(profile), tuple(shape), str(device))
    pipe = _PIPELINES.get(key)
    if pipe is None:
        if len(_PIPELINES) > 8:
            _PIPELINES.clear()
        pipe = PoissonPipeline(profile, shape, device)
        _PIPELINES[key] = pipe
    return pipe


def _gaussian_pipeline(profile, shape, device, resolved) -> GaussianPipeline:
    key = ("gaussian", repr(profile), tuple(shape), str(device))
    pipe = _PIPELINES.get(key)
    if pipe is None:
        if len(_PIPELINES) > 8:
            _PIPELINES.clear()
        pipe = GaussianPipeline(profile, shape, device, resolved)
        _PIPELINES[key] = pipe
    return pipe


def _validated_input(volume, profile, sigma, sigma_psd, max_cache_bytes):
    """The reference's argument checks, in its order (bmndalgo.py:124-155):
    returns the float32 input and the resolved profile."""
    if sigma is not None and sigma_psd is not None:
        raise ValueError("sigma and sigma_psd are mutually exclusive.")
    if sigma is not None:
        _validate_sigma(sigma)
    _validate_max_cache_bytes(max_cache_bytes)
    resolved = resolve_profile(profile, tuple(int(s) for s in np.shape(volume)))
    if resolved.is_poisson and sigma_psd is not None:
        raise ValueError("sigma_psd is only supported for stationary additive noise; "
                         "direct Poisson mode uses a variance volume instead.")
    arr = np.asarray(volume)
    if not np.all(np.isfinite(arr)):
        raise ValueError("volume must contain only finite values.")
    with np.errstate(over="ignore", invalid="ignore"):
        vol = arr.astype(np.float32, copy=False)
    if not np.all(np.isfinite(vol)):
        raise ValueError("volume values must be representable as finite float32 values.")
    if resolved.is_poisson and np.any(vol < 0.0):
        raise ValueError("Direct Poisson input must contain only non-negative values.")
    if sigma_psd is not None:
        validate_sigma_psd(tuple(int(s) for s in vol.shape), sigma_psd)
    return vol, resolved


def bmnd(volume, profile, sigma: float | None = None, sigma_psd=None,
         return_sigma_map: bool = False,
         stage_callback: Callable[[str, np.ndarray], None] | None = None,
         blockmatch_callback: Callable[[str, BlockMatchResult], None] | None = None,
         max_cache_bytes=CacheMode.AUTO):
    """Denoise an N-D Poisson or stationary-Gaussian volume (2-D and 3-D on
    the MI355X path); for Gaussian noise `sigma` or `sigma_psd` gives the
    noise level, else it is estimated from the volume.

    Returns a float32 array of the input shape or, with ``return_sigma_map``,
    ``(array, {"global": g, "ht": map, "wiener": map})`` as the reference
    does (bmndalgo.py:1048-1062; maps of zeros and no "global" entry for an
    all-zero input).  Callbacks receive copies of
    the HT pilot / Wiener estimate (``stage_callback``) and the block-match
    results (``blockmatch_callback``), in the reference's order."""
    vol, resolved = _validated_input(volume, profile, sigma, sigma_psd, max_cache_bytes)
    if not np.any(vol != 0.0):
        if stage_callback is not None:
            stage_callback("ht", vol.copy())
            stage_callback("wiener", vol.copy())
        if return_sigma_map:
            zeros = np.zeros_like(vol, dtype=np.float32)
            return vol, {"ht": zeros, "wiener": zeros}
        return vol
    check_supported(profile, resolved)
    vol = np.ascontiguousarray(vol)
    device = _device()
    if not resolved.is_poisson:
        return _bmnd_gaussian(vol, profile, resolved, device, sigma, sigma_psd,
                              return_sigma_map, stage_callback, blockmatch_callback)
    pipe = _pipeline(profile, vol.shape, device)
    if (USE_GRAPHS and vol.size <= GRAPH_MAX_VOXELS and stage_callback is None
            and blockmatch_callback is None and not return_sigma_map and pipe.graph_supported()):
        return pipe.run_graph(torch.from_numpy(vol)).cpu().numpy()
    vol_t = torch.from_numpy(vol).to(device)

    def match_hook(stage: str):
        # blockmatch_callback(stage, result) right after that stage's matching
        # and before its filtering, as the reference calls it
        # (bmndalgo.py:325-326, 742-743)
        if blockmatch_callback is None:
            return None
        plan = pipe.plans[stage]
        counts = tuple(vol.shape[d] - plan.block[d] + 1 for d in range(vol.ndim))

        def hook(sel, cnt, acc):
            blockmatch_callback(stage, build_result(plan.schedule, counts, sel.cpu().numpy(),
                                                    cnt.cpu().numpy(), acc.cpu().numpy()))
        return hook

    ht, var1, global_sigma = pipe.run_ht(vol_t, sigma_maps=return_sigma_map,
                                         on_match=match_hook("ht"))
    if stage_callback is not None:
        # the pilot, before stage 2 starts (bmndalgo.py:632-633)
        stage_callback("ht", ht.est.cpu().numpy())
    wi = pipe.run_wiener(vol_t, ht.est, var1, global_sigma=global_sigma,
                         on_match=match_hook("wiener"))
    del var1
    out = wi.est.cpu().numpy()
    if stage_callback is not None:
        stage_callback("wiener", out.copy())
    if return_sigma_map:
        # bmndalgo.py:1048-1062: per-stage maps, global sigma where no patch
        # reaches a voxel
        return out, {"global": global_sigma, "ht": ht.sigma_map.cpu().numpy(),
                     "wiener": wi.sigma_map.cpu().numpy()}
    return out


def _match_hook(pipe, shape, blockmatch_callback):
    """stage -> hook(sel, counts, accepted) calling blockmatch_callback(stage,
    result) after that stage's matching (bmndalgo.py:325-326, 742-743)."""
    def make(stage: str):
        if blockmatch_callback is None:
            return None
        plan = pipe.plans[stage]
        counts = tuple(shape[d] - plan.block[d] + 1 for d in range(len(shape)))

        def hook(sel, cnt, acc):
            blockmatch_callback(stage, build_result(plan.schedule, counts, sel.cpu().numpy(),
                                                    cnt.cpu().numpy(), acc.cpu().numpy()))
        return hook
    return make


def _bmnd_gaussian(vol, profile, resolved, device, sigma, sigma_psd, return_sigma_map,
                   stage_callback, blockmatch_callback):
    """The stationary Gaussian branches of bmndalgo.py:203-1063 on the device."""
    pipe = _gaussian_pipeline(profile, vol.shape, device, resolved)
    vol_t = torch.from_numpy(vol).to(device)

    def on_pilot(pilot):
        if stage_callback is not None:
            stage_callback("ht", pilot.cpu().numpy())
    ht, wi, global_sigma = pipe.run(vol_t, sigma=sigma, sigma_psd=sigma_psd,
                                    sigma_maps=return_sigma_map,
                                    on_match=_match_hook(pipe, vol.shape, blockmatch_callback),
                                    on_pilot=on_pilot, vol_host=vol)
    out = wi.est.cpu().numpy()
    if stage_callback is not None:
        stage_callback("wiener", out.copy())
    if return_sigma_map:
        return out, {"global": global_sigma, "ht": ht.sigma_map.cpu().numpy(),
                     "wiener": wi.sigma_map.cpu().numpy()}
    return out


def bmnd_batch(volumes, profile, sigma: float | None = None,
               max_cache_bytes=CacheMode.AUTO) -> list[np.ndarray]:
    """Denoise several inputs of one dimensionality and one profile, of any
    shapes, as one batch: every kernel runs once over the references (or
    aggregation tiles) of all inputs together, which fills the GPU with
    inputs too small to fill it alone.  No counterpart in the reference; each
    input is checked as bmnd() checks it and its result is bmnd(input,
    profile)'s, bit for bit.  Under the Gaussian model the noise tables
    depend on each input's shape (and, with sigma None, on its estimated
    sigma), so the inputs run one after another through the Gaussian
    pipeline.  Returns float32 arrays in input order."""
    volumes = list(volumes)
    if len({np.ndim(v) for v in volumes}) > 1:
        raise ValueError("every input of a batch must have the same number of dimensions.")
    items = [_validated_input(v, profile, sigma, None, max_cache_bytes) for v in volumes]
    outs: list = [None] * len(items)
    live = []
    for i, (vol, resolved) in enumerate(items):
        if not np.any(vol != 0.0):
            outs[i] = vol  # the reference's all-zero early return
        else:
            check_supported(profile, resolved)
            live.append(i)
    if live and not items[live[0]][1].is_poisson:
        device = _device()
        for i in live:
            vol, resolved = items[i]
            vol = np.ascontiguousarray(vol)
            pipe = _gaussian_pipeline(profile, vol.shape, device, resolved)
            _, wi, _ = pipe.run(torch.from_numpy(vol).to(device), sigma=sigma, vol_host=vol)
            outs[i] = wi.est.cpu().numpy()
        return outs
    if live:
        device = torch.device("cuda", torch.cuda.current_device())
        probe = _pipeline(profile, items[live[0]][0].shape, device)
        if any(probe.mass.values()) or any(probe.exact_risk.values()) or probe.ssd:
            # group-mass conservation, the exact group risk and the SSD
            # distance run input by input
            for i in live:
                vol = np.ascontiguousarray(items[i][0])
                pipe = _pipeline(profile, vol.shape, device)
                _, wi = pipe.run(torch.from_numpy(vol).to(device))
                outs[i] = wi.est.cpu().numpy()
            return outs
        live = [live[j] for j in table_first_order([items[i][0] for i in live])]
        pipe = BatchPipeline(profile, [items[i][0].shape for i in live], device)
        _, wi = pipe.run(pipe.pack([items[i][0] for i in live]))
        for j, res in enumerate(pipe.unpack(wi.est)):
            outs[live[j]] = res
    return outs
