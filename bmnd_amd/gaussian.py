"""Device pipeline of the stationary Gaussian bmnd(): both stages on one GPU.

Per stage: SSD block matching (stage 1 with the noise-bias correction
gamma * noise SSD per lag unless nf asked for the conventional path) ->
exact-plane group noise variance from the cross-covariance maps -> group
filtering (Gaussian hard threshold / Wiener gain) -> ordered aggregation.
The noise description -- the PSD, its nf processing, the per-lag noise SSD
and the per-coefficient cross-covariance maps -- is a set of small host
tables per (profile, sigma) built by gaussianvariance with the reference's
numpy steps and uploaded once; every per-group and per-voxel step runs in
the HIP kernels.  Mirrors bmndalgo.py:71-1063 on the Gaussian branches.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .engine import (
    StageOutput,
    StagePlan,
    _host3,
    _NoTimer,
    _ptr,
    aggregate_stage,
    filter_code,
    make_plan,
)
from .enums import BlockMatchDistance
from .gaussianvariance import (
    GaussianVarianceModel,
    blockmatch_noise_ssd,
    build_variance_model,
    cross_covariance_maps,
)
from .profiles import BMNDProfile
from .profilevalidation import resolve_profile
from .shifts import LazySchedule
from .transforms import spatial_matrices

# filter-code bit of the stationary Gaussian model (bmnd_hip_filter_groups)
GAUSSIAN_CODE = 256


@dataclass
class GaussianTables:
    """One noise description on the device: the model, each stage's
    lag-major cross-covariance maps [D][P] and the stage-1 noise SSD."""
    model: GaussianVarianceModel
    domain3: tuple[int, int, int]
    cov: dict
    noise_ssd: torch.Tensor | None


class GaussianPipeline:
    """Two-stage stationary-Gaussian denoiser on one device; buffers are
    torch tensors on `device`, kernels run on the current torch stream."""

    def __init__(self, profile: BMNDProfile, shape: tuple[int, ...], device=None, resolved=None):
        self.profile = profile
        self.shape = tuple(int(s) for s in shape)
        self.ndim = len(self.shape)
        self.shape3 = _host3(self.shape, self.ndim, 1)
        self.device = torch.device(device if device is not None else "cuda")
        resolved = resolved if resolved is not None else resolve_profile(profile, self.shape)
        if resolved.is_poisson:
            raise ValueError("GaussianPipeline runs the stationary Gaussian noise model")
        if resolved.blockmatch_distance is not BlockMatchDistance.SSD:
            raise NotImplementedError("the Gaussian model runs with the 'ssd' block-matching "
                                      "distance on the MI355X path.")
        # bmndalgo.py:246-247, profilevalidation.py:445-448: k = 0 and no
        # gamma correction on the conventional nf path
        self.k = int(resolved.variance_k)
        self.nf = resolved.effective_nf
        self.conventional = bool(resolved.conventional_nf)
        self.gamma = 0.0 if self.conventional else float(profile.gamma)
        self.plans = {s: make_plan(profile, s, self.ndim) for s in ("ht", "wiener")}
        for plan in self.plans.values():
            counts = tuple(self.shape[d] - plan.block[d] + 1 for d in range(self.ndim))
            plan.lazy_schedule = LazySchedule(
                counts, plan.block, plan.step, mode=profile.reference_schedule_mode,
                shift_density=profile.reference_shift_density,
                schedule_density=profile.reference_schedule_density)
        self.filter_codes = {"ht": filter_code(profile, "ht", 0) | GAUSSIAN_CODE,
                             "wiener": filter_code(profile, "wiener", 1) | GAUSSIAN_CODE}
        self._refs: dict = {}
        self._tables: dict = {}

    # -- noise description ---------------------------------------------------------
    def estimate_sigma(self, vol: torch.Tensor, vol_host: np.ndarray | None = None) -> float:
        """estimate_gaussian_global_sigma (sigmaestimation.py:7-45) of the
        device tensor `vol`: the MAD / 0.6745 of its finite nonzero values,
        median and MAD found on the device (bmnd_hip_median_mad).  When the
        MAD is 0 the reference falls back to the IQR, a trimmed std and
        sqrt(eps); those rare cases are evaluated on the host copy."""
        out = torch.empty(2, dtype=torch.float32, device=self.device)
        cnt = torch.empty(2, dtype=torch.int64, device=self.device)
        _lib.check(_lib.lib().bmnd_hip_median_mad(_ptr(vol), vol.numel(), _ptr(out), _ptr(cnt),
                                                  self._stream()))
        mad = float(out.cpu().numpy()[1])
        sigma = mad / 0.6745
        if sigma > 0.0 and np.isfinite(sigma):
            return sigma
        from .psd import estimate_gaussian_global_sigma
        return estimate_gaussian_global_sigma(vol_host if vol_host is not None
                                              else vol.cpu().numpy())

    def tables(self, vol: torch.Tensor | None, sigma=None, sigma_psd=None,
               vol_host: np.ndarray | None = None) -> GaussianTables:
        """The noise description of (sigma, sigma_psd), built on first use:
        with neither given, sigma is the MAD estimate of `vol`
        (variance.py:461-462)."""
        if sigma is None and sigma_psd is None:
            sigma = self.estimate_sigma(vol, vol_host)
        if sigma_psd is not None:
            arr = np.asarray(sigma_psd, dtype=np.float32)
            key = ("psd", arr.shape, arr.tobytes())
        else:
            key = ("sigma", float(sigma))
        t = self._tables.get(key)
        if t is not None:
            return t
        if len(self._tables) > 4:
            self._tables.clear()
        model = build_variance_model(np.zeros(self.shape, np.float32), sigma=sigma,
                                     sigma_psd=sigma_psd, nf=self.nf)
        domain = tuple(int(d) for d in model.variance_domain_shape)
        cov = {}
        for name, plan in self.plans.items():
            mats = [spatial_matrices(plan.spatial, b)[0] for b in plan.block]
            maps = cross_covariance_maps(model, plan.block, mats)
            lag_major = np.ascontiguousarray(maps.reshape(maps.shape[0], -1).T)
            cov[name] = torch.from_numpy(lag_major).to(self.device)
        noise = None
        if self.gamma != 0.0:
            plan = self.plans["ht"]
            radius = plan.radius3[3 - self.ndim:]
            ssd = blockmatch_noise_ssd(model, plan.block, radius)
            noise = torch.from_numpy(np.ascontiguousarray(ssd.reshape(-1))).to(self.device)
        t = GaussianTables(model, _host3(domain, self.ndim, 1), cov, noise)
        self._tables[key] = t
        return t

    # -- helpers ---------------------------------------------------------------------
    def _stream(self) -> ctypes.c_void_p:
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _ref_lin(self, plan: StagePlan) -> torch.Tensor:
        if plan.name not in self._refs:
            self._refs[plan.name] = plan.lazy_schedule.device_lin(self.device)
        return self._refs[plan.name]

    def ssd_threshold(self, name: str, vol: torch.Tensor) -> float:
        """The stage's SSD threshold: the profile's times the squared value
        range of the matched volume, as float32 (blockmatching.py:168-184,
        1446-1449)."""
        rng = torch.empty(2, dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().bmnd_hip_value_range(_ptr(vol), vol.numel(), _ptr(rng),
                                                   self._stream()))
        lo, hi = (np.float32(x) for x in rng.cpu().numpy())
        scale = float(hi - lo)
        if scale <= 0.0:
            scale = 1.0
        thr = float(getattr(self.profile, f"{name}_match_threshold"))
        return float(np.float32(thr * scale * scale))

    def stage(self, name: str, match_vol: torch.Tensor, noisy: torch.Tensor,
              tables: GaussianTables, timer=None, sigma_fill: float | None = None,
              pilot: torch.Tensor | None = None, on_match=None) -> StageOutput:
        """One stage (bmndalgo.py:286-633 for "ht", 648-1046 for "wiener")."""
        L = _lib.lib()
        timer = timer or _NoTimer()
        plan = self.plans[name]
        stream = self._stream()
        shape = (ctypes.c_int64 * 3)(*self.shape3)
        block = (ctypes.c_int32 * 3)(*plan.block3)
        radius = (ctypes.c_int32 * 3)(*plan.radius3)
        domain = (ctypes.c_int32 * 3)(*tables.domain3)
        refs = self._ref_lin(plan)
        R = refs.numel()
        dev = self.device
        sel = torch.empty(R * plan.max_m, dtype=torch.int32, device=dev)
        cnt = torch.empty(R, dtype=torch.int32, device=dev)
        acc = torch.empty(R, dtype=torch.int32, device=dev)
        thr = self.ssd_threshold(name, match_vol)
        gamma = self.gamma if name == "ht" else 0.0
        noise = tables.noise_ssd if gamma != 0.0 else None
        tok = timer.start(f"{name}.blockmatch")
        _lib.check(L.bmnd_hip_blockmatch_ssd(_ptr(match_vol), shape, block, radius, _ptr(refs), R,
                                             thr, plan.max_m, plan.min_m, 1, _ptr(noise), gamma,
                                             _ptr(sel), _ptr(cnt), _ptr(acc), stream))
        timer.stop(tok)
        if on_match is not None:
            on_match(sel.view(R, plan.max_m), cnt, acc)
        mats, kai = plan.device_constants(dev)
        cov = tables.cov[name]
        kcap = min(self.k, plan.max_m)
        exact = None
        if kcap > 0:
            exact = torch.empty(R * kcap * plan.P, dtype=torch.float32, device=dev)
            tok = timer.start(f"{name}.gaussian_variance")
            _lib.check(L.bmnd_hip_gaussian_variance(_ptr(cov), domain, shape, block, radius,
                                                    _ptr(sel), _ptr(cnt), R, plan.max_m,
                                                    _ptr(mats), self.k, _ptr(exact), stream))
            timer.stop(tok)
        filt = torch.empty(R * plan.max_m * plan.P, dtype=torch.float32, device=dev)
        wts = torch.empty(R * plan.max_m, dtype=torch.float32, device=dev)
        wiener = name == "wiener"
        gsig = None
        if sigma_fill is not None:
            gsig = torch.empty(R if wiener else R * plan.max_m * plan.P, dtype=torch.float32,
                               device=dev)
        tok = timer.start(f"{name}.filter")
        _lib.check(L.bmnd_hip_filter_groups_sigma(self.filter_codes[name], _ptr(noisy),
                                                  _ptr(pilot if wiener else None), _ptr(cov),
                                                  shape, block, radius, _ptr(sel), _ptr(cnt), R,
                                                  plan.max_m, _ptr(mats), _ptr(exact),
                                                  plan.shrink, self.k, 1.0, _ptr(filt),
                                                  _ptr(wts), _ptr(gsig), stream))
        timer.stop(tok)
        del exact
        tok = timer.start(f"{name}.aggregate")
        est, smap = aggregate_stage(plan, self.shape3, refs, R, filt, wts, sel, cnt, noisy,
                                    noisy, gsig, wiener, sigma_fill, stream)
        timer.stop(tok)
        del filt, wts, gsig
        return StageOutput(est, sel.view(R, plan.max_m), cnt, acc, plan.lazy_schedule, smap)

    def run(self, vol: torch.Tensor, sigma=None, sigma_psd=None, timer=None,
            sigma_maps: bool = False, on_match=None, on_pilot=None, vol_host=None):
        """Both stages on the float32 contiguous device tensor `vol`; returns
        (HT output, Wiener output, global sigma).  on_match(stage) gives the
        stage's match hook or None; on_pilot(pilot) runs between the stages;
        `vol_host`, the same volume on the host, spares a copy when sigma is
        estimated."""
        if vol.dtype != torch.float32 or not vol.is_contiguous() or tuple(vol.shape) != self.shape:
            raise ValueError("vol must be a contiguous float32 tensor of the planned shape")
        t = timer or _NoTimer()
        tok = t.start("gaussian.tables")
        tables = self.tables(vol, sigma, sigma_psd, vol_host)
        t.stop(tok)
        gsig = tables.model.global_sigma
        fill = float(np.float32(gsig)) if sigma_maps else None
        hook = on_match or (lambda stage: None)
        ht = self.stage("ht", vol, vol, tables, timer, fill, on_match=hook("ht"))
        if on_pilot is not None:
            on_pilot(ht.est)
        wi = self.stage("wiener", ht.est, vol, tables, timer, fill, pilot=ht.est,
                        on_match=hook("wiener"))
        return ht, wi, gsig
