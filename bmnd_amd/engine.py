"""Device pipeline of the Poisson bmnd(): both stages on one GPU.

Per stage: schedule (built on the device, kept per shape) -> fused block matching ->
group filtering -> ordered aggregation, all on the caller's HIP stream over
torch-allocated device buffers.  2-D inputs run as (1, H, W) volumes.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .aggregationweights import constant_synthesis_directions, synthesis_metric_diagonals
from .enums import (
    AggregationWeightDomain,
    AggregationWeightModel,
    AggregationWeightScope,
    BlockMatchDistance,
    NoiseModel,
    PoissonGroupMassConservation,
    PoissonHTMatchPolicy,
    PoissonVarianceSource,
    PoissonWienerGainMode,
    coerce_enum,
)
from .blockmatching import (
    compute_poisson_reference_thresholds,
    poisson_null_moment_table_f32,
    resolve_blockmatch_distance,
)
from .profiles import BMNDProfile
from .shifts import LazySchedule, ScheduleArrays
from .transforms import Transform, TransformMode, TransformType, group_matrices, spatial_matrices
from .utils import kaiser_window_nd

# Summation order of each stage's exact overlap-aware planes:
#   "serial" -- bmnd_hip_exact_std: the reference's order (union voxels in C
#               order, members increasing inside a voxel), every product and
#               sum a separate float32 operation; bit for bit the oracle's.
#   "pairs"  -- bmnd_hip_exact_std_pairs: the square expanded over member
#               pairs, each pair's overlap sum a separable transform, the plane
#               products on the matrix cores; the same quantity to float32
#               rounding, several times faster.
# Stage 1's planes decide hard-threshold masks and through them stage 2's
# lists, so stage 1 keeps the serial order; the Wiener planes enter the gains
# and weights continuously and take the pair form.
EXACT_ORDERS = ("serial", "pairs")
EXACT_ORDER = {"ht": "serial", "wiener": "pairs"}
K_PAIRS_FLAG = {"ht": 1 << 16, "wiener": 1 << 17}  # BMND_HIP_EXACT_PAIRS_HT / _WIENER
# The Wiener filter's member-axis products (the group transforms' n x n
# matrix products): "mfma" on the matrix cores (v_mfma_f32_16x16x4f32, the
# products summed in the hardware's order: the same values to float32
# rounding) or "serial" (the reference's fmaf chains, bit for bit the
# oracle's).  Stage 1 always runs the serial form: its coefficients meet the
# hard threshold.
TRANSFORMS = ("serial", "mfma")
WIENER_TRANSFORMS = ["mfma"]
K_MFMA_WIENER = 1 << 18       # BMND_HIP_MFMA_WIENER
FILTER_MFMA = 1024            # filter code bit 10


def set_exact_order(ht: str | None = None, wiener: str | None = None,
                    transforms: str | None = None) -> dict:
    """Set the exact planes' summation order per stage ("serial" or "pairs")
    and the Wiener filter's member transforms ("serial" or "mfma") for the
    calls that follow; returns the previous setting."""
    prev = dict(EXACT_ORDER, transforms=WIENER_TRANSFORMS[0])
    for name, value in (("ht", ht), ("wiener", wiener)):
        if value is None:
            continue
        if value not in EXACT_ORDERS:
            raise ValueError(f"exact order must be one of {EXACT_ORDERS}, got {value!r}")
        EXACT_ORDER[name] = value
    if transforms is not None:
        if transforms not in TRANSFORMS:
            raise ValueError(f"transforms must be one of {TRANSFORMS}, got {transforms!r}")
        WIENER_TRANSFORMS[0] = transforms
    return prev


def k_with_order(k: int) -> int:
    """k with the pair-order and matrix-core flags of the current setting
    OR'd in (the k of bmnd_hip_stage and bmnd_hip_pipeline_desc)."""
    k = int(k) | sum(K_PAIRS_FLAG[st] for st in ("ht", "wiener") if EXACT_ORDER[st] == "pairs")
    return k | (K_MFMA_WIENER if WIENER_TRANSFORMS[0] == "mfma" else 0)


def code_with_transforms(code: int, name: str) -> int:
    """A Wiener stage's filter code with the matrix-core bit of the current
    setting."""
    if name == "wiener" and WIENER_TRANSFORMS[0] == "mfma":
        return int(code) | FILTER_MFMA
    return int(code)


def exact_std_entry(name: str, batch: bool = False):
    """The library entry point computing stage `name`'s exact planes in the
    current order."""
    L = _lib.lib()
    pairs = EXACT_ORDER[name] == "pairs"
    if batch:
        return L.bmnd_hip_exact_std_pairs_batch if pairs else L.bmnd_hip_exact_std_batch
    return L.bmnd_hip_exact_std_pairs if pairs else L.bmnd_hip_exact_std


# (block, search radius) pairs compiled into libbmnd_hip.so, in 3-D form
SUPPORTED = {
    ((1, 8, 8), (0, 19, 19)),
    ((4, 4, 4), (7, 7, 7)),
    ((5, 5, 5), (7, 7, 7)),
}


class KernelTimer:
    """Records HIP events around each library call on the current stream so
    per-kernel durations can be read after a synchronize."""

    def __init__(self):
        self.spans: dict[str, list] = {}

    def start(self, name: str):
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        return (name, ev)

    def stop(self, token) -> None:
        name, ev0 = token
        ev1 = torch.cuda.Event(enable_timing=True)
        ev1.record()
        self.spans.setdefault(name, []).append((ev0, ev1))

    def totals_ms(self) -> dict[str, list[float]]:
        return {k: [a.elapsed_time(b) for a, b in v] for k, v in self.spans.items()}


class _NoTimer:
    def start(self, name):
        return None

    def stop(self, token):
        return None


def _ptr(t: torch.Tensor | None) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _host3(values, ndim: int, pad: int) -> tuple[int, int, int]:
    v = tuple(int(x) for x in values)
    return (pad,) * (3 - ndim) + v


def _spatial_kind(transform, stage: str) -> TransformType:
    items = [transform] if isinstance(transform, Transform) else list(transform)
    kinds = {t.mode: t.type for t in items}
    if (len(items) == 2 and kinds.get(TransformMode.GROUP) is TransformType.HAAR
            and kinds.get(TransformMode.SPATIAL) in (TransformType.BIOR15, TransformType.DCT)):
        return kinds[TransformMode.SPATIAL]
    raise NotImplementedError(
        f"{stage}_transform must be [HAAR/GROUP, BIOR15 or DCT/SPATIAL] on the MI355X path.")


@dataclass
class StagePlan:
    name: str
    ndim: int
    block: tuple[int, ...]
    step: tuple[int, ...]
    block3: tuple[int, int, int]
    radius3: tuple[int, int, int]
    max_m: int
    min_m: int
    threshold: float
    shrink: float
    spatial: TransformType
    beta: float
    kaiser: np.ndarray
    lazy_schedule: LazySchedule | None = None
    _dev: dict = field(default_factory=dict)

    @property
    def schedule(self) -> ScheduleArrays:
        """The host arrays of the stage's schedule (built on first use)."""
        return self.lazy_schedule.arrays

    @property
    def P(self) -> int:
        return int(np.prod(self.block3))

    def matrix_table(self) -> np.ndarray:
        parts = []
        for n in range(1, self.max_m + 1):
            parts.append(group_matrices(n)[0].reshape(-1))
        for n in range(1, self.max_m + 1):
            parts.append(group_matrices(n)[1].reshape(-1))
        fwd, inv = [], []
        for b in self.block3:
            if b == 1:
                f = i = np.ones((1, 1), dtype=np.float32)
            else:
                f, i = spatial_matrices(self.spatial, b)
            fwd.append(f.reshape(-1))
            inv.append(i.reshape(-1))
        table = np.concatenate(parts + fwd + inv).astype(np.float32)
        # float64 part from the next even offset: the synthesis metric
        # diagonals of the windowed-synthesis weights and the constant-synthesis
        # directions of the group-mass correction
        ginv = [group_matrices(n)[1] for n in range(1, self.max_m + 1)]
        sinv = [i.reshape(b, b) for i, b in zip(inv, self.block3)]
        smd, gmd = synthesis_metric_diagonals(sinv, self.kaiser, ginv)
        mdir = np.concatenate(constant_synthesis_directions(ginv))
        sdir = np.concatenate(constant_synthesis_directions(sinv))
        f64 = np.concatenate([smd, gmd, mdir, sdir]).astype(np.float64)
        if table.size % 2:
            table = np.concatenate([table, np.zeros(1, np.float32)])
        # float32 pair tables of the pair-order exact planes: per axis and
        # offset e, H[t][p] = A[t][p] A[t][p - e + b - 1] (0 off the block)
        pair_tabs = []
        for f, b in zip(fwd, self.block3):
            A = f.reshape(b, b).astype(np.float32)
            h = np.zeros((2 * b - 1, b, b), np.float32)
            for e in range(2 * b - 1):
                for p in range(b):
                    q = p - (e - (b - 1))
                    if 0 <= q < b:
                        h[e, :, p] = A[:, p] * A[:, q]
            pair_tabs.append(h.reshape(-1))
        # the spatial metric SM = S^T diag(w^2) S of the exact group risk
        # (variance.py:1543-1556), float32 from float64 sums
        S = np.ones((1, 1))
        for i in sinv:
            S = np.kron(S, np.asarray(i, np.float64))
        w2 = (np.asarray(self.kaiser, np.float32).reshape(-1) ** 2).astype(np.float64)
        sm = (S.T @ (w2[:, None] * S)).astype(np.float32)
        table = np.concatenate([table, f64.view(np.float32)] + pair_tabs + [sm.reshape(-1)])
        blk = (ctypes.c_int32 * 3)(*self.block3)
        assert table.size == _lib.lib().bmnd_hip_matrix_table_size(self.max_m, blk)
        return table

    def device_constants(self, device) -> tuple[torch.Tensor, torch.Tensor]:
        key = str(device)
        if key not in self._dev:
            mats = torch.from_numpy(self.matrix_table()).to(device)
            kai = torch.from_numpy(np.ascontiguousarray(self.kaiser.reshape(-1))).to(device)
            self._dev[key] = (mats, kai)
        return self._dev[key]


# reference distance codes (blockmatching.py:13-26) of the Poisson distances
_DISTANCE_CODES = {BlockMatchDistance.POISSON_DEVIANCE: 1, BlockMatchDistance.PEARSON: 2,
                   BlockMatchDistance.ANSCOMBE_SSD: 3}


# Wiener gain signal power (wiener.py:24-53) -> the filters' wiener argument
_GAIN_ARG = {PoissonWienerGainMode.VARIANCE_SCALED: 1, PoissonWienerGainMode.CLASSIC: 2,
             PoissonWienerGainMode.NOISE_FLOOR: 3}


def _ht_policy(profile: BMNDProfile) -> PoissonHTMatchPolicy:
    return coerce_enum(profile.poisson_ht_match_policy, PoissonHTMatchPolicy,
                       "poisson_ht_match_policy")


def check_supported(profile: BMNDProfile, resolved) -> None:
    """Raise NotImplementedError for profile options outside the MI355X path."""
    if not resolved.is_poisson:
        if resolved.blockmatch_distance is not BlockMatchDistance.SSD:
            raise NotImplementedError("the Gaussian model runs with the 'ssd' block-matching "
                                      "distance on the MI355X path.")
    elif resolved.blockmatch_distance not in _DISTANCE_CODES \
            and resolved.blockmatch_distance is not BlockMatchDistance.SSD:
        raise NotImplementedError("the 'poisson_deviance', 'pearson', 'anscombe_ssd' and 'ssd' "
                                  "block-matching distances run on the MI355X path.")
    def field(name, enum):
        return coerce_enum(getattr(profile, name), enum, name)

    if not resolved.is_poisson:
        for stage in ("ht", "wiener"):
            if exact_risk_stage(profile, stage, int(resolved.variance_k)):
                raise NotImplementedError(
                    f"the exact covariance-aware {stage} group risk (windowed-synthesis domain, "
                    "group scope, k > 0) runs under the Poisson model on the MI355X path.")


def exact_risk_stage(profile: BMNDProfile, stage: str, k: int) -> bool:
    """Whether stage `stage` weights its groups by the exact covariance-aware
    group risk (bmndalgo.py:359-377, 752-758; wiener.py:638-844): the
    variance model (or, in the Wiener stage, the risk model) in the
    windowed-synthesis domain with group scope and k > 0."""
    def field(name, enum):
        return coerce_enum(getattr(profile, name), enum, name)
    models = ((AggregationWeightModel.VARIANCE,) if stage == "ht" else
              (AggregationWeightModel.VARIANCE, AggregationWeightModel.RISK))
    return (field(f"{stage}_weight_model", AggregationWeightModel) in models
            and field(f"{stage}_weight_domain", AggregationWeightDomain)
            is AggregationWeightDomain.WINDOWED_SYNTHESIS
            and field(f"{stage}_weight_scope", AggregationWeightScope)
            is AggregationWeightScope.GROUP and int(k) > 0)


# the filters' weight-model codes (aggregationweights.py:19-21 order differs:
# 0 here is the default variance model)
_WEIGHT_MODEL_CODE = {AggregationWeightModel.VARIANCE: 0, AggregationWeightModel.CLASSIC: 1,
                      AggregationWeightModel.RISK: 2}


def filter_code(profile: BMNDProfile, stage: str, gain_arg: int, coefficients: bool = False) -> int:
    """The filter code of bmnd_hip_filter_groups for one stage: the stage
    and Wiener gain in bits 0-1, soft thresholding (bit 2), the weight model
    (bits 3-4), the windowed-synthesis domain (bit 5), the group scope (bit 6)
    and, for the group-mass correction, coefficient output (bit 7)."""
    def field(name, enum):
        return coerce_enum(getattr(profile, name), enum, name)
    code = gain_arg if stage == "wiener" else 0
    if stage == "ht" and profile.ht_use_soft_thresholding:
        code |= 4
    code |= _WEIGHT_MODEL_CODE[field(f"{stage}_weight_model", AggregationWeightModel)] << 3
    if field(f"{stage}_weight_domain", AggregationWeightDomain) \
            is AggregationWeightDomain.WINDOWED_SYNTHESIS:
        code |= 32
    if field(f"{stage}_weight_scope", AggregationWeightScope) is AggregationWeightScope.GROUP:
        code |= 64
    if coefficients:
        code |= 128
    return code


def make_plan(profile: BMNDProfile, stage: str, ndim: int) -> StagePlan:
    if ndim not in (2, 3):
        raise NotImplementedError("the MI355X path runs 2-D images and 3-D volumes.")
    block = tuple(int(b) for b in getattr(profile, f"{stage}_block_size"))
    radius = tuple(int(r) for r in getattr(profile, f"{stage}_search_window"))
    block3, radius3 = _host3(block, ndim, 1), _host3(radius, ndim, 0)
    if (block3, radius3) not in SUPPORTED:
        raise NotImplementedError(
            f"{stage} block {block} with search radius {radius} is not compiled into the "
            "MI355X kernels (supported: 8x8/19x19, 4^3/7^3, 5^3/7^3).")
    max_m = int(getattr(profile, f"{stage}_max_stack_size"))
    if max_m > (16 if block3 == (4, 4, 4) else 32):
        raise NotImplementedError(f"{stage}_max_stack_size={max_m} exceeds the compiled group size.")
    beta = float(getattr(profile, f"{stage}_kaiser_beta"))
    return StagePlan(
        name=stage, ndim=ndim, block=block, step=tuple(int(s) for s in getattr(profile, f"{stage}_step")),
        block3=block3, radius3=radius3, max_m=max_m,
        min_m=int(getattr(profile, f"{stage}_min_stack_size")),
        threshold=float(np.float32(getattr(profile, f"{stage}_match_threshold"))),
        shrink=float(profile.ht_lambda_threshold if stage == "ht" else profile.wiener_variance_scale),
        spatial=_spatial_kind(getattr(profile, f"{stage}_transform"), stage),
        beta=beta, kaiser=kaiser_window_nd(block, beta))


@dataclass
class SurrogateStages:
    pilot: torch.Tensor       # the stage-1 pilot stage 2 ran on
    wiener: "StageOutput"


@dataclass
class StageOutput:
    est: torch.Tensor
    sel: torch.Tensor
    counts: torch.Tensor
    accepted: torch.Tensor
    schedule: LazySchedule
    sigma_map: torch.Tensor | None = None  # with sigma maps requested


def blockmatch_work(plan: StagePlan, shape: tuple[int, ...],
                    schedule: ScheduleArrays | None = None) -> tuple[int, int]:
    """(candidate pairs, deviance terms) of a stage: sum over references of
    the clipped search-box size, times the patch volume for terms (the
    plan's schedule unless `schedule` is given, as for a batch input)."""
    ndim = len(shape)
    counts = np.array([shape[d] - plan.block[d] + 1 for d in range(ndim)], dtype=np.int64)
    radius = np.array(plan.radius3[3 - ndim:], dtype=np.int64)
    ref = (schedule if schedule is not None else plan.schedule).reference_abs
    lo = np.maximum(ref - radius, 0)
    hi = np.minimum(ref + radius, counts - 1)
    pairs = int(np.prod(hi - lo + 1, axis=1).sum())
    return pairs, pairs * int(np.prod(plan.block))


def exact_std_work(plan: StagePlan, shape3: tuple[int, int, int], sel: torch.Tensor,
                   counts: torch.Tensor, k: int, chunk: int = 16384) -> dict:
    """Algorithmic work of a stage's exact-variance launch, counted on the
    device: for every group of n members, P = patch size, U = voxels covered
    by the union of its patches and q = min(k, n) exact planes, the
    overlap-aware variance sum_v var_v (sum_(i,p) G[q,i] S[s,p])^2 costs
    q*P*(2*n*P + 3*U) floating-point operations (a multiply and an add per
    entry and coefficient, two multiplies and an add per union voxel);
    `flops_mac` counts the multiply-adds alone, 2*q*P*n*P per group.
    `pairs` counts the member pairs i <= j whose patches overlap, the work
    items of the pair order: each costs three separable passes over whole
    rows, 2 * P * (b0 + b1 + b2) flops on the vector units (`flops_pairs`),
    and 2 * 4 * P flops of plane products on the matrix cores
    (`flops_pairs_mfma`, four planes per product)."""
    b0, b1, b2 = plan.block3
    P = b0 * b1 * b2
    n0, n1, n2 = shape3
    c1, c2 = n1 - b1 + 1, n2 - b2 + 1
    dev = sel.device
    max_m = plan.max_m
    sel2 = sel.view(-1, max_m).to(torch.int64)
    cnt = counts.to(torch.int64)
    off = torch.arange(P, device=dev)
    p2, p1, p0 = off % b2, (off // b2) % b1, off // (b2 * b1)
    poff = (p0 * n1 + p1) * n2 + p2  # voxel offset of each patch element
    entries = 0
    union = 0
    flops = 0
    flops_pairs = 0
    npairs = 0
    blk = torch.tensor(plan.block3, device=dev)
    for a in range(0, sel2.shape[0], chunk):
        s = sel2[a:a + chunk]
        n = cnt[a:a + chunk]
        x2, x1, x0 = s % c2, (s // c2) % c1, s // (c2 * c1)
        base = (x0 * n1 + x1) * n2 + x2  # voxel index of each member's origin
        vox = base[:, :, None] + poff[None, None, :]
        valid = (torch.arange(max_m, device=dev)[None, :] < n[:, None])[:, :, None]
        vox = torch.where(valid, vox, torch.full_like(vox, -1)).view(s.shape[0], -1)
        vs, _ = torch.sort(vox, dim=1)
        new = torch.ones_like(vs, dtype=torch.bool)
        new[:, 1:] = vs[:, 1:] != vs[:, :-1]
        u = (new & (vs >= 0)).sum(dim=1)
        q = torch.clamp(n, max=k)
        entries += int((n * P).sum())
        union += int(u.sum())
        flops += int((q * P * (2 * n * P + 3 * u)).sum())
        flops_pairs += int((2 * q * P * n * P).sum())
        # overlapping member pairs i <= j
        pos = torch.stack([x0, x1, x2], dim=-1)
        d = (pos[:, :, None, :] - pos[:, None, :, :]).abs()
        ov = (d < blk).all(dim=-1)
        live = torch.arange(max_m, device=dev)[None, :] < n[:, None]
        ov = ov & live[:, :, None] & live[:, None, :]
        npairs += int((ov.sum() + ov.diagonal(dim1=1, dim2=2).sum()) // 2)
    return {"entries": entries, "union_voxels": union, "flops": flops,
            "flops_mac": flops_pairs, "pairs": npairs,
            "flops_pairs": npairs * 2 * P * (b0 + b1 + b2),
            "flops_pairs_mfma": npairs * 2 * 4 * P}


def aggregate_stage(plan: StagePlan, shape3, refs: torch.Tensor, R: int, filt: torch.Tensor,
                    wts: torch.Tensor, sel: torch.Tensor, cnt: torch.Tensor, noisy: torch.Tensor,
                    fallback: torch.Tensor, gsig: torch.Tensor | None, wiener: bool,
                    sigma_fill: float | None, stream) -> tuple[torch.Tensor, torch.Tensor | None]:
    """Ordered aggregation of a stage's filtered groups (patchaggregation.py:
    170-199, bmndalgo.py:626-631, 1039-1044): the estimate, `fallback` where
    no patch lands, and with `gsig` the sigma map (`sigma_fill` where no
    patch lands)."""
    L = _lib.lib()
    shape = (ctypes.c_int64 * 3)(*shape3)
    block = (ctypes.c_int32 * 3)(*plan.block3)
    radius = (ctypes.c_int32 * 3)(*plan.radius3)
    _, kai = plan.device_constants(noisy.device)
    npatch = int(np.prod([shape3[d] - plan.block3[d] + 1 for d in range(3)]))
    ref_index = torch.empty(npatch, dtype=torch.int32, device=noisy.device)
    _lib.check(L.bmnd_hip_reference_map(_ptr(refs), R, _ptr(ref_index), npatch, stream))
    est = torch.empty_like(noisy)
    smap = torch.empty_like(noisy) if gsig is not None else None
    _lib.check(L.bmnd_hip_aggregate_sigma(_ptr(filt), _ptr(wts), _ptr(sel), _ptr(cnt),
                                          _ptr(ref_index), R, plan.max_m, shape, block, radius,
                                          _ptr(kai), _ptr(fallback), _ptr(est), _ptr(gsig),
                                          int(wiener), float(sigma_fill or 0.0), _ptr(smap),
                                          stream))
    return est, smap


def _threshold_options(profile: BMNDProfile) -> tuple:
    """The options compute_poisson_reference_thresholds takes after the
    schedule and block, from the profile (bmndalgo.py:273-283)."""
    return (float(profile.poisson_count_scale), float(profile.ht_match_threshold),
            float(profile.poisson_ht_match_structure_beta),
            float(profile.poisson_ht_match_intensity_power),
            int(profile.poisson_ht_match_moment_max_count))


class PoissonPipeline:
    """Two-stage Poisson denoiser on one device.  Buffers are torch tensors on
    `device`; kernels run on the current torch stream."""

    DEVIANCE_MODES = ("auto", "exact", "one-log", "auto-one-log", "certified", "auto-certified",
                      "fast", "auto-fast")

    def __init__(self, profile: BMNDProfile, shape: tuple[int, ...], device=None,
                 deviance: str = "auto-fast"):
        self.profile = profile
        self.shape = tuple(int(s) for s in shape)
        self.ndim = len(self.shape)
        self.shape3 = _host3(self.shape, self.ndim, 1)
        self.device = torch.device(device if device is not None else "cuda")
        if deviance not in self.DEVIANCE_MODES:
            raise ValueError(f"deviance must be one of {self.DEVIANCE_MODES}, got {deviance!r}")
        # Integer counts <= 100 always use the float64 term table (the reference
        # expression, evaluated on the host with the C library's log).  Other
        # inputs: "auto" evaluates the reference's two-log expression term by
        # term; "auto-certified" evaluates the one-log form
        # r log 2r + c log 2c - s log s with a degree-4 log together with a
        # bound on its distance from the reference expression, and re-evaluates
        # the reference expression for the rare candidates whose tie-grid key
        # the bound does not pin, so its keys equal those of "auto";
        # "auto-fast" (default) bounds every candidate's key with a float32
        # pass first and runs the certified evaluation only for the
        # candidates that can change the selection or the accepted count,
        # with the same result;
        # "auto-one-log" uses the one-log form alone, which can move a
        # distance by ~1e-14 and so a key across the 32-ulp tie grid in rare
        # cases.
        self.deviance = deviance
        self.last_modes: dict[str, int] = {}
        self.count_scale = float(profile.poisson_count_scale)
        self.floor = float(profile.poisson_variance_floor)
        self.distance = resolve_blockmatch_distance(profile.blockmatch_distance, NoiseModel.POISSON)
        # candidate-standardized stage-1 distance (bmndalgo.py:311-316)
        self.standardized = _ht_policy(profile) is PoissonHTMatchPolicy.CANDIDATE_STANDARDIZED
        # the SSD distance: thresholds scaled by the matched volume's squared
        # value range, stage 1 with the Poisson gamma correction
        # (bmndalgo.py:285-294, 321-322)
        self.ssd = self.distance is BlockMatchDistance.SSD
        self.gamma = float(profile.gamma) if self.ssd else 0.0
        # the filters' stage argument: 0 hard threshold, 1-3 the Wiener gain
        self.wiener_arg = _GAIN_ARG[coerce_enum(profile.poisson_wiener_gain_mode,
                                                PoissonWienerGainMode, "poisson_wiener_gain_mode")]
        # group-mass conservation per stage (bmndalgo.py:263-266, 704-711)
        mass = coerce_enum(profile.poisson_group_mass_conservation, PoissonGroupMassConservation,
                           "poisson_group_mass_conservation")
        self.mass = {"ht": mass in (PoissonGroupMassConservation.HT, PoissonGroupMassConservation.BOTH),
                     "wiener": mass in (PoissonGroupMassConservation.WIENER,
                                        PoissonGroupMassConservation.BOTH)}
        # each stage's full filter code (bmnd_hip_filter_groups)
        self.filter_codes = {st: filter_code(profile, st, self.wiener_arg, self.mass[st])
                             for st in ("ht", "wiener")}
        self._null = None
        self.k = int(profile.k)
        # stages weighted by the exact covariance-aware group risk
        self.exact_risk = {st: exact_risk_stage(profile, st, self.k) for st in ("ht", "wiener")}
        self.plans = {s: make_plan(profile, s, self.ndim) for s in ("ht", "wiener")}
        for plan in self.plans.values():
            counts = tuple(self.shape[d] - plan.block[d] + 1 for d in range(self.ndim))
            plan.lazy_schedule = LazySchedule(
                counts, plan.block, plan.step, mode=profile.reference_schedule_mode,
                shift_density=profile.reference_shift_density,
                schedule_density=profile.reference_schedule_density)
        self._refs = {}

    # -- helpers -----------------------------------------------------------------
    def _wiener_sources(self, vol: torch.Tensor, pilot: torch.Tensor, var1: torch.Tensor, t):
        """Volume the Wiener stage matches on and its variance volume: the
        pilot and the pilot's variance by default, the observation and the
        stage-1 variance with the OBSERVATION sources (bmndalgo.py:635-646,
        714-720)."""
        def source(field):
            return coerce_enum(getattr(self.profile, field), PoissonVarianceSource, field)
        on_obs = source("poisson_wiener_match_source") is PoissonVarianceSource.OBSERVATION
        if source("poisson_variance_source_wiener") is PoissonVarianceSource.OBSERVATION:
            var2 = var1
        else:
            tok = t.start("wiener.variance")
            var2 = self.variance(pilot)
            t.stop(tok)
        return (vol if on_obs else pilot), var2

    def _stream(self) -> ctypes.c_void_p:
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _ref_lin(self, plan: StagePlan) -> torch.Tensor:
        """The stage's references (patch-grid linear indices), built on the
        device on first use and kept."""
        if plan.name not in self._refs:
            self._refs[plan.name] = plan.lazy_schedule.device_lin(self.device)
        return self._refs[plan.name]

    def null_moments(self) -> tuple[torch.Tensor, int]:
        """Device null table of the candidate-standardized distance and its
        max pooled count."""
        if self._null is None:
            mc = self.profile.poisson_ht_match_moment_max_count
            tab = poisson_null_moment_table_f32(self.distance, mc)
            self._null = (torch.from_numpy(tab).to(self.device), int(mc))
        return self._null

    def blockmatch_mode(self, vol: torch.Tensor, name: str = "wiener") -> tuple[int, int]:
        """(kernel mode, table max) for matching stage `name` on `vol`: the
        deviance modes below, 5 / 6 for the Pearson / Anscombe SSD distances,
        7-9 for the standardized stage-1 distances, 10 for SSD and 11 for
        SSD with the Poisson gamma correction (stage 1, gamma != 0)."""
        if self.ssd:
            return (11 if name == "ht" and self.gamma != 0.0 else 10), 0
        code = _DISTANCE_CODES[self.distance]
        if name == "ht" and self.standardized:
            return 6 + code, 0
        if code != 1:
            return 3 + code, 0
        if self.deviance == "exact":
            return 0, 0
        if self.deviance == "one-log":
            return 1, 0
        if self.deviance == "certified":
            return 3, 0
        if self.deviance == "fast":
            return 4, 0
        fallback = {"auto": 0, "auto-one-log": 1, "auto-certified": 3,
                    "auto-fast": 4}[self.deviance]
        if self.count_scale != 1.0:
            return fallback, 0
        flags = torch.empty(2, dtype=torch.int32, device=self.device)
        _lib.check(_lib.lib().bmnd_hip_count_range(_ptr(vol), vol.numel(), _ptr(flags),
                                                   self._stream()))
        is_int, vmax = (int(x) for x in flags.cpu().tolist())
        if is_int and vmax <= 100:
            return 2, vmax
        return fallback, 0

    def ssd_threshold(self, name: str, vol: torch.Tensor) -> float:
        """The stage's SSD threshold: the profile's times the squared value
        range of the matched volume, as float32 (blockmatching.py:168-184)."""
        rng = torch.empty(2, dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().bmnd_hip_value_range(_ptr(vol), vol.numel(), _ptr(rng),
                                                   self._stream()))
        lo, hi = (np.float32(x) for x in rng.cpu().numpy())
        scale = float(hi - lo)
        if scale <= 0.0:
            scale = 1.0
        thr = float(getattr(self.profile, f"{name}_match_threshold"))
        return float(np.float32(thr * scale * scale))

    def variance(self, vol: torch.Tensor) -> torch.Tensor:
        out = torch.empty_like(vol)
        _lib.check(_lib.lib().bmnd_hip_variance_volume(_ptr(vol), _ptr(out), vol.numel(),
                                                       self.count_scale, self.floor, self._stream()))
        return out

    def ht_thresholds(self, vol: torch.Tensor, timer=None) -> torch.Tensor | None:
        """Per-reference stage-1 thresholds of the "reference_finite_count"
        policy (bmndalgo.py:269-284), formed on the host from the input's
        reference patches and uploaded; None under the fixed policy."""
        if _ht_policy(self.profile) is not PoissonHTMatchPolicy.REFERENCE_FINITE_COUNT:
            return None
        t = timer or _NoTimer()
        tok = t.start("ht.thresholds")
        plan = self.plans["ht"]
        thr = compute_poisson_reference_thresholds(
            vol.cpu().numpy().reshape(self.shape), plan.schedule, plan.block,
            *_threshold_options(self.profile), distance=self.distance)
        out = torch.from_numpy(thr).to(self.device)
        t.stop(tok)
        return out

    def global_sigma(self, var: torch.Tensor) -> float:
        """Global noise std of the Poisson variance model of `var` (the
        stage-1 variance volume): the sigma maps' "global" value."""
        g = ctypes.c_double(0.0)
        _lib.check(_lib.lib().bmnd_hip_poisson_global_sigma(_ptr(var), var.numel(), self.floor,
                                                            ctypes.byref(g), self._stream()))
        return float(g.value)

    def stage(self, name: str, match_vol: torch.Tensor, noisy: torch.Tensor, var: torch.Tensor,
              fallback: torch.Tensor, timer=None, sigma_fill: float | None = None,
              pilot: torch.Tensor | None = None,
              ref_thresholds: torch.Tensor | None = None, on_match=None) -> StageOutput:
        """One stage.  With `sigma_fill` (the global sigma) the stage also
        builds its sigma map, `sigma_fill` where no patch reaches a voxel.
        The Wiener gains use `pilot` (default: the matched volume).
        `ref_thresholds` (float32, one per reference) replaces the plan's
        threshold.  `on_match(sel, counts, accepted)` runs after the block
        matching and before the filtering (the point where the reference
        calls its blockmatch_callback, bmndalgo.py:325-326, 742-743)."""
        L = _lib.lib()
        timer = timer or _NoTimer()
        plan = self.plans[name]
        stream = self._stream()
        shape = (ctypes.c_int64 * 3)(*self.shape3)
        block = (ctypes.c_int32 * 3)(*plan.block3)
        radius = (ctypes.c_int32 * 3)(*plan.radius3)
        refs = self._ref_lin(plan)
        R = refs.numel()
        dev = self.device
        sel = torch.empty(R * plan.max_m, dtype=torch.int32, device=dev)
        cnt = torch.empty(R, dtype=torch.int32, device=dev)
        acc = torch.empty(R, dtype=torch.int32, device=dev)
        mode, tmax = self.blockmatch_mode(match_vol, name)
        self.last_modes[name] = mode
        tok = timer.start(f"{name}.blockmatch")
        if mode == 11:
            _lib.check(L.bmnd_hip_blockmatch_ssd_poisson(
                _ptr(match_vol), shape, block, radius, _ptr(refs), R,
                self.ssd_threshold(name, match_vol), plan.max_m, plan.min_m, 1,
                self.gamma, self.count_scale, _ptr(sel), _ptr(cnt), _ptr(acc), stream))
        elif mode == 10:
            _lib.check(L.bmnd_hip_blockmatch_ssd(
                _ptr(match_vol), shape, block, radius, _ptr(refs), R,
                self.ssd_threshold(name, match_vol), plan.max_m, plan.min_m, 1, None, 0.0,
                _ptr(sel), _ptr(cnt), _ptr(acc), stream))
        elif mode >= 7:
            null, null_max = self.null_moments()
            _lib.check(L.bmnd_hip_blockmatch_standardized(
                _ptr(match_vol), shape, block, radius, _ptr(refs), R, plan.threshold, plan.max_m,
                plan.min_m, 1, 1, self.count_scale, mode - 6, _ptr(null), null_max, _ptr(sel),
                _ptr(cnt), _ptr(acc), stream))
        elif ref_thresholds is None:
            _lib.check(L.bmnd_hip_blockmatch(_ptr(match_vol), shape, block, radius, _ptr(refs), R,
                                             plan.threshold, plan.max_m, plan.min_m, 1, 1,
                                             self.count_scale, mode, tmax, _ptr(sel), _ptr(cnt),
                                             _ptr(acc), stream))
        else:
            if ref_thresholds.numel() != R or ref_thresholds.dtype != torch.float32:
                raise ValueError(f"ref_thresholds must hold {R} float32 values")
            _lib.check(L.bmnd_hip_blockmatch_thresholds(
                _ptr(match_vol), shape, block, radius, _ptr(refs), _ptr(ref_thresholds), R,
                plan.max_m, plan.min_m, 1, 1, self.count_scale, mode, tmax, _ptr(sel), _ptr(cnt),
                _ptr(acc), stream))
        timer.stop(tok)
        if on_match is not None:
            on_match(sel.view(R, plan.max_m), cnt, acc)
        mats, kai = plan.device_constants(dev)
        kcap = min(self.k, plan.max_m)
        exact = None
        if kcap > 0:
            exact = torch.empty(R * kcap * plan.P, dtype=torch.float32, device=dev)
            tok = timer.start(f"{name}.exact_std")
            _lib.check(exact_std_entry(name)(_ptr(var), shape, block, radius, _ptr(sel),
                                             _ptr(cnt), R, plan.max_m, _ptr(mats), self.k,
                                             _ptr(exact), stream))
            timer.stop(tok)
        filt = torch.empty(R * plan.max_m * plan.P, dtype=torch.float32, device=dev)
        wts = torch.empty(R * plan.max_m, dtype=torch.float32, device=dev)
        wiener = name == "wiener"
        gsig = None
        if sigma_fill is not None:
            # Wiener: one std per group; HT: one per member coefficient
            gsig = torch.empty(R if wiener else R * plan.max_m * plan.P, dtype=torch.float32,
                               device=dev)
        tok = timer.start(f"{name}.filter")
        filter_args = (code_with_transforms(self.filter_codes[name], name), _ptr(noisy),
                       _ptr((match_vol if pilot is None else pilot) if wiener else None),
                       _ptr(var), shape, block, radius, _ptr(sel), _ptr(cnt), R, plan.max_m,
                       _ptr(mats), _ptr(exact), plan.shrink, self.k, self.count_scale,
                       _ptr(filt), _ptr(wts), _ptr(gsig))
        if self.exact_risk[name]:
            # the exact covariance-aware group risk replaces the weights
            need = L.bmnd_hip_exact_risk_workspace(R, plan.max_m, self.k, block)
            ws = torch.empty(max(need, 1), dtype=torch.uint8, device=dev)
            _lib.check(L.bmnd_hip_filter_groups_exact(*filter_args, _ptr(ws), need, stream))
            del ws
        else:
            _lib.check(L.bmnd_hip_filter_groups_sigma(*filter_args, stream))
        timer.stop(tok)
        del exact
        npatch = int(np.prod([self.shape3[d] - plan.block3[d] + 1 for d in range(3)]))
        ref_index = torch.empty(npatch, dtype=torch.int32, device=dev)
        _lib.check(L.bmnd_hip_reference_map(_ptr(refs), R, _ptr(ref_index), npatch, stream))
        est = torch.empty_like(noisy)
        smap = torch.empty_like(noisy) if gsig is not None else None
        tok = timer.start(f"{name}.aggregate")
        if self.mass[name]:
            # group-mass conservation (poisson.py:235-265): the denominator
            # (and sigma map) of the plain aggregation, every group's
            # coefficients corrected against it and synthesised, then the
            # numerator over the unchanged denominator without eps
            den = torch.empty_like(noisy)
            _lib.check(L.bmnd_hip_aggregate_norm(2, _ptr(filt), _ptr(wts), _ptr(sel), _ptr(cnt),
                                                 _ptr(ref_index), R, plan.max_m, shape, block,
                                                 radius, _ptr(kai), _ptr(fallback), _ptr(den),
                                                 _ptr(gsig), int(wiener),
                                                 float(sigma_fill or 0.0), _ptr(smap), stream))
            _lib.check(L.bmnd_hip_mass_correct(_ptr(filt), _ptr(wts), _ptr(sel), _ptr(cnt), R,
                                               plan.max_m, shape, block, radius, _ptr(den),
                                               _ptr(noisy), _ptr(kai), _ptr(mats), stream))
            _lib.check(L.bmnd_hip_aggregate_norm(1, _ptr(filt), _ptr(wts), _ptr(sel), _ptr(cnt),
                                                 _ptr(ref_index), R, plan.max_m, shape, block,
                                                 radius, _ptr(kai), _ptr(fallback), _ptr(est),
                                                 None, 0, 0.0, None, stream))
            del den
        else:
            _lib.check(L.bmnd_hip_aggregate_sigma(_ptr(filt), _ptr(wts), _ptr(sel), _ptr(cnt),
                                                  _ptr(ref_index), R, plan.max_m, shape, block,
                                                  radius, _ptr(kai), _ptr(fallback), _ptr(est),
                                                  _ptr(gsig), int(wiener),
                                                  float(sigma_fill or 0.0), _ptr(smap), stream))
        timer.stop(tok)
        del filt, wts, ref_index, gsig
        return StageOutput(est, sel.view(R, plan.max_m), cnt, acc, plan.lazy_schedule, smap)

    def run_ht(self, vol: torch.Tensor, timer=None, sigma_maps: bool = False, on_match=None):
        """Stage 1 (bmndalgo.py:213-633) on the float32 contiguous device
        tensor `vol` of `shape`: returns the HT stage output (its `est` is the
        pilot), the stage-1 variance volume and, with `sigma_maps`, the
        global sigma (else None)."""
        if vol.dtype != torch.float32 or not vol.is_contiguous() or tuple(vol.shape) != self.shape:
            raise ValueError("vol must be a contiguous float32 tensor of the planned shape")
        t = timer or _NoTimer()
        tok = t.start("ht.variance")
        var1 = self.variance(vol)
        t.stop(tok)
        gsig = self.global_sigma(var1) if sigma_maps else None
        fill = None if gsig is None else float(np.float32(gsig))
        thr = self.ht_thresholds(vol, t)
        ht = self.stage("ht", vol, vol, var1, vol, timer, fill, ref_thresholds=thr,
                        on_match=on_match)
        return ht, var1, gsig

    def run_wiener(self, vol: torch.Tensor, pilot: torch.Tensor, var1: torch.Tensor, timer=None,
                   global_sigma: float | None = None, on_match=None) -> StageOutput:
        """Stage 2 (bmndalgo.py:635-1046) from the observation `vol`, the
        stage-1 pilot and the stage-1 variance volume; with `global_sigma`
        the stage also builds its sigma map."""
        t = timer or _NoTimer()
        fill = None if global_sigma is None else float(np.float32(global_sigma))
        match2, var2 = self._wiener_sources(vol, pilot, var1, t)
        return self.stage("wiener", match2, vol, var2, vol, timer, fill, pilot=pilot,
                          on_match=on_match)

    def run_surrogate(self, vol: torch.Tensor, net, check, timer=None, on_pilot=None):
        """Stage 2 on a network pilot (bmnd_amd.surrogate): the network's
        pilot of `vol`, its agreement check against the plain stage 1 on
        sub-volumes, the plain stage 1 on `vol` instead when it fails, then
        the reference's stage 2 on the pilot kept; on_pilot(pilot) runs
        between the two (the place of the reference's "ht" callbacks).
        Returns (SurrogateStages, SurrogateInfo)."""
        from .surrogate import SurrogateInfo, anscombe_stats
        t = timer or _NoTimer()
        # the count level first: E[A(x)^2] / 4 - 3/8 = E[x] for A(x) =
        # 2 sqrt(x + 3/8); above the network's range it is not run at all
        tok = t.start("surrogate.check")
        a_mean, a_std = anscombe_stats(vol)
        mean_count = (a_std * a_std + a_mean * a_mean) / 4.0 - 0.375
        t.stop(tok)
        pilot, dist = None, float("nan")
        if mean_count <= check.max_mean_count:
            tok = t.start("surrogate.network")
            pilot = net.pilot(vol)
            t.stop(tok)
            tok = t.start("surrogate.check")
            dist = check.distance(self.profile, vol, pilot)
            t.stop(tok)
        if pilot is None or not check.accepts(dist, mean_count):
            ht, var1, _ = self.run_ht(vol, timer)
            pilot, used = ht.est, "plain"
        else:
            tok = t.start("ht.variance")
            var1 = self.variance(vol)
            t.stop(tok)
            used = "network"
        if on_pilot is not None:
            on_pilot(pilot)
        wi = self.run_wiener(vol, pilot, var1, timer)
        return SurrogateStages(pilot, wi), SurrogateInfo(used, dist, check.threshold, mean_count)

    # ---- both stages as one HIP graph ------------------------------------------
    def graph_supported(self) -> bool:
        """Whether the pipeline's options run in a captured graph: the fixed
        stage-1 threshold and the plain (not standardized) distances; sigma
        maps, callbacks and the non-default thresholding and weights are served
        by the call-by-call path."""
        return (_ht_policy(self.profile) is PoissonHTMatchPolicy.FIXED and not self.standardized
                and not self.ssd
                and self.filter_codes == {"ht": 0, "wiener": self.wiener_arg})

    def _graph_modes(self, vol: torch.Tensor) -> tuple[int, int, int]:
        """(stage-1 mode, its table size, stage-2 mode) for a capture: stage 1
        as the call-by-call path picks it for `vol`; stage 2 the mode that
        path takes for a non-integer pilot (every mode it could take returns
        the same lists)."""
        m1, tmax = self.blockmatch_mode(vol, "ht")
        return m1, tmax, self._graph_mode_wiener()

    def _graph_mode_wiener(self) -> int:
        code = _DISTANCE_CODES[self.distance]
        if code != 1:
            return 3 + code
        return {"auto": 0, "exact": 0, "one-log": 1, "auto-one-log": 1, "certified": 3,
                "auto-certified": 3, "fast": 4, "auto-fast": 4}[self.deviance]

    def run_graph(self, vol) -> torch.Tensor:
        """Both stages by one launch of the captured graph (StageGraph) on
        `vol` (a float32 array or device tensor of `shape`); returns the
        Wiener estimate (a buffer of the graph, overwritten by its next
        launch).  The graph is captured on the first call, and again if an
        input breaks the count-table precondition it was captured with
        (that launch is then repeated with the new graph)."""
        if not self.graph_supported():
            raise NotImplementedError("this profile's options run on the call-by-call path")
        g = getattr(self, "_graph", None)
        if g is None:
            self._graph = g = StageGraph(self)
        g.input.copy_(torch.as_tensor(vol).view(self.shape), non_blocking=False)
        if g.handle is None or g.k != k_with_order(self.k):
            g.capture(*self._graph_modes(g.input))
        rc = g.launch()
        if rc == _lib_einval():
            g.capture(*self._graph_modes(g.input))
            rc = g.launch()
        _lib.check(rc)
        return g.output

    def run(self, vol: torch.Tensor, timer=None, sigma_maps: bool = False):
        """vol: float32 contiguous device tensor of `shape`.  Returns the HT
        (pilot) and Wiener stage outputs and, with `sigma_maps`, the global
        sigma as a third value (each stage output then carries its map)."""
        ht, var1, gsig = self.run_ht(vol, timer, sigma_maps)
        wi = self.run_wiener(vol, ht.est, var1, timer, gsig)
        del var1
        if sigma_maps:
            return ht, wi, gsig
        return ht, wi


def _lib_einval() -> int:
    return -1  # BMND_HIP_EINVAL


def _pipeline_desc(pipe: PoissonPipeline, plans: dict, refs: dict, lists: dict, modes,
                   vin: torch.Tensor, pilot: torch.Tensor, out: torch.Tensor):
    """bmnd_hip_pipeline_desc for `pipe`'s options with each stage's plan,
    reference list (device int32) and match-list buffers (sel, counts,
    accepted); the caller sets the geometry (shape, or the batch fields)."""
    d = _lib.PipelineDesc()
    for st, name in enumerate(("ht", "wiener")):
        plan = plans[name]
        d.block[st][:] = plan.block3
        d.radius[st][:] = plan.radius3
        d.ref_lin[st] = refs[name].data_ptr()
        d.nrefs[st] = refs[name].numel()
        d.threshold[st] = plan.threshold
        d.max_matches[st], d.min_matches[st] = plan.max_m, plan.min_m
        mats, kai = plan.device_constants(pipe.device)
        d.matrices[st], d.kaiser[st] = mats.data_ptr(), kai.data_ptr()
        d.shrink[st] = plan.shrink
        sel, cnt, acc = lists[name]
        d.sel[st], d.counts[st], d.accepted[st] = sel.data_ptr(), cnt.data_ptr(), acc.data_ptr()
    d.mode[0], d.table_max, d.mode[1] = modes
    d.k, d.wiener_gain = k_with_order(pipe.k), pipe.wiener_arg
    d.count_scale, d.variance_floor = pipe.count_scale, pipe.floor
    src = lambda f: coerce_enum(getattr(pipe.profile, f), PoissonVarianceSource, f)  # noqa: E731
    d.match_observation = int(src("poisson_wiener_match_source")
                              is PoissonVarianceSource.OBSERVATION)
    d.variance_observation = int(src("poisson_variance_source_wiener")
                                 is PoissonVarianceSource.OBSERVATION)
    d.input, d.pilot, d.output = vin.data_ptr(), pilot.data_ptr(), out.data_ptr()
    return d


class StageGraph:
    """Both stages of a PoissonPipeline captured as one HIP graph
    (bmnd_hip_pipeline_*): the graph reads `input` and writes `pilot`,
    `output` and each stage's match lists at every launch."""

    def __init__(self, pipe: PoissonPipeline):
        self.pipe = pipe
        dev = pipe.device
        self.input = torch.empty(pipe.shape, dtype=torch.float32, device=dev)
        self.pilot = torch.empty_like(self.input)
        self.output = torch.empty_like(self.input)
        self.lists = {}
        for name, plan in pipe.plans.items():
            R = pipe._ref_lin(plan).numel()
            self.lists[name] = (torch.empty(R * plan.max_m, dtype=torch.int32, device=dev),
                                torch.empty(R, dtype=torch.int32, device=dev),
                                torch.empty(R, dtype=torch.int32, device=dev))
        self.handle = None
        self.modes = None
        self.k = None

    def capture(self, mode_ht: int, table_max: int, mode_wi: int) -> None:
        self.release()
        pipe, L = self.pipe, _lib.lib()
        refs = {name: pipe._ref_lin(plan) for name, plan in pipe.plans.items()}
        d = _pipeline_desc(pipe, pipe.plans, refs, self.lists, (mode_ht, table_max, mode_wi),
                           self.input, self.pilot, self.output)
        d.shape[:] = pipe.shape3
        h = ctypes.c_void_p()
        torch.cuda.synchronize(pipe.device)
        _lib.check(L.bmnd_hip_pipeline_create(ctypes.byref(d), ctypes.byref(h)))
        self.handle, self.modes = h, (mode_ht, table_max, mode_wi)
        self.k = d.k

    def launch(self) -> int:
        """One launch on the current stream, then its status (synchronises)."""
        L = _lib.lib()
        stream = self.pipe._stream()
        rc = L.bmnd_hip_pipeline_launch(self.handle, stream)
        if rc:
            return rc
        return L.bmnd_hip_pipeline_status(self.handle, stream)

    def release(self) -> None:
        if self.handle is not None:
            torch.cuda.synchronize(self.pipe.device)
            _lib.lib().bmnd_hip_pipeline_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


# bmnd_hip_input (include/bmnd_hip.h): one input of a batch
INPUT_DTYPE = np.dtype([("vox_off", "<i8"), ("grid_off", "<i8"), ("tile_off", "<i8"),
                        ("shape", "<i4", (3,)), ("grid", "<i4", (3,))])
assert INPUT_DTYPE.itemsize == 48


@dataclass
class BatchStage:
    """One stage's view of a batch: per-input schedules, the flat reference
    list (patch-grid index within the input, and the input), and the input
    records on the host and on the device."""
    plan: StagePlan
    schedules: list
    ref_lin: torch.Tensor
    ref_input: torch.Tensor
    ref_off: np.ndarray        # first flat reference of every input, plus the total
    inputs_h: np.ndarray       # INPUT_DTYPE records
    inputs_d: torch.Tensor     # the same bytes on the device
    grid_total: int

    @property
    def nrefs(self) -> int:
        return int(self.ref_off[-1])

    def args(self):
        """(inputs_h, inputs_d, ninputs) as the C-ABI takes them."""
        return (ctypes.c_void_p(self.inputs_h.ctypes.data), _ptr(self.inputs_d),
                len(self.inputs_h))


@dataclass
class BatchStageOutput:
    est: torch.Tensor       # packed estimates
    sel: torch.Tensor       # (nrefs, max_m), patch indices within each input
    counts: torch.Tensor
    accepted: torch.Tensor
    stage: BatchStage


def table_first_order(arrays) -> list[int]:
    """Batch order that puts inputs of integer counts <= 100 first (stable):
    stage-1 matching then takes them in one launch of the count-table kernel
    and the others in one launch of the float-pass kernel."""
    def table(v):
        v = np.asarray(v)
        return bool(np.all(v == np.rint(v)) and float(v.max()) <= 100.0)
    return sorted(range(len(arrays)), key=lambda i: not table(arrays[i]))


class BatchPipeline:
    """The two-stage Poisson denoiser over a batch of inputs of one profile
    and one dimensionality but any shapes: each kernel runs once over the
    flat list of every input's references (or aggregation tiles), reading the
    inputs from one packed device buffer.  Every input's result is the one
    PoissonPipeline gives for it alone."""

    def __init__(self, profile: BMNDProfile, shapes, device=None, deviance: str = "auto-fast"):
        shapes = [tuple(int(s) for s in sh) for sh in shapes]
        if not shapes:
            raise ValueError("a batch needs at least one input")
        ndims = {len(s) for s in shapes}
        if len(ndims) != 1:
            raise ValueError("every input of a batch must have the same number of dimensions")
        self.ndim = ndims.pop()
        self.shapes = shapes
        self.shape3s = [_host3(s, self.ndim, 1) for s in shapes]
        self.vox = np.array([int(np.prod(s)) for s in self.shape3s], dtype=np.int64)
        self.vox_off = np.concatenate([[0], np.cumsum(self.vox)]).astype(np.int64)
        self.device = torch.device(device if device is not None else "cuda")
        # matching-mode choice and the per-call constants are PoissonPipeline's
        self._single = PoissonPipeline(profile, shapes[0], self.device, deviance)
        if any(self._single.mass.values()):
            raise NotImplementedError("group-mass conservation runs input by input (bmnd()); "
                                      "bmnd_batch does not take it.")
        if any(self._single.exact_risk.values()):
            raise NotImplementedError("the exact covariance-aware group risk runs input by input "
                                      "(bmnd()); bmnd_batch does not take it.")
        if self._single.ssd:
            raise NotImplementedError("the SSD distance under the Poisson model runs input by "
                                      "input (bmnd()); BatchPipeline does not take it.")
        self.stages = {name: self._stage_view(profile, name) for name in ("ht", "wiener")}
        self.last_runs: dict[str, list] = {}  # (mode, first input, end) per matching launch

    @property
    def total_voxels(self) -> int:
        return int(self.vox_off[-1])

    def _stage_view(self, profile: BMNDProfile, name: str) -> BatchStage:
        plan = make_plan(profile, name, self.ndim)
        tile = (ctypes.c_int32 * 3)()
        _lib.check(_lib.lib().bmnd_hip_aggregate_tile((ctypes.c_int32 * 3)(*plan.block3), tile))
        recs = np.zeros(len(self.shapes), dtype=INPUT_DTYPE)
        schedules, lins, inputs = [], [], []
        grid_off = tile_off = 0
        ref_off = [0]
        dev = self.device
        for i, (shape, shape3) in enumerate(zip(self.shapes, self.shape3s)):
            counts = tuple(shape[d] - plan.block[d] + 1 for d in range(self.ndim))
            sched = LazySchedule(
                counts, plan.block, plan.step, mode=profile.reference_schedule_mode,
                shift_density=profile.reference_shift_density,
                schedule_density=profile.reference_schedule_density)
            schedules.append(sched)
            lin = sched.device_lin(dev)
            lins.append(lin)
            inputs.append(torch.full((lin.numel(),), i, dtype=torch.int32, device=dev))
            grid3 = tuple(shape3[d] - plan.block3[d] + 1 for d in range(3))
            recs[i] = (self.vox_off[i], grid_off, tile_off, shape3, grid3)
            grid_off += int(np.prod(grid3))
            tile_off += int(np.prod([-(-shape3[d] // tile[d]) for d in range(3)]))
            ref_off.append(ref_off[-1] + lin.numel())
        return BatchStage(
            plan=plan, schedules=schedules,
            ref_lin=torch.cat(lins), ref_input=torch.cat(inputs),
            ref_off=np.asarray(ref_off, dtype=np.int64), inputs_h=recs,
            inputs_d=torch.from_numpy(recs.view(np.uint8).copy()).to(dev),
            grid_total=grid_off)

    def pack(self, volumes) -> torch.Tensor:
        """One float32 device buffer holding every input in order."""
        flat = np.concatenate([np.ascontiguousarray(v, dtype=np.float32).reshape(-1)
                               for v in volumes])
        return torch.from_numpy(flat).to(self.device)

    def unpack(self, packed: torch.Tensor) -> list[np.ndarray]:
        host = packed.cpu().numpy()
        return [host[self.vox_off[i]:self.vox_off[i + 1]].reshape(self.shapes[i]).copy()
                for i in range(len(self.shapes))]

    def input_modes(self, vol: torch.Tensor, name: str = "wiener") -> list[tuple[int, int]]:
        """(matching mode, table max) of every input of the packed `vol`, as
        PoissonPipeline.blockmatch_mode chooses them for the input alone."""
        single = self._single
        n = len(self.shapes)
        code = _DISTANCE_CODES[single.distance]
        if code != 1 or (name == "ht" and single.standardized):
            return [single.blockmatch_mode(vol, name)] * n
        fixed = {"exact": 0, "one-log": 1, "certified": 3, "fast": 4}
        if single.deviance in fixed:
            return [(fixed[single.deviance], 0)] * n
        fallback = {"auto": 0, "auto-one-log": 1, "auto-certified": 3,
                    "auto-fast": 4}[single.deviance]
        if single.count_scale != 1.0:
            return [(fallback, 0)] * n
        L = _lib.lib()
        flags = torch.empty(2 * n, dtype=torch.int32, device=self.device)
        base, fbase = vol.data_ptr(), flags.data_ptr()
        for i in range(n):
            _lib.check(L.bmnd_hip_count_range(ctypes.c_void_p(base + 4 * int(self.vox_off[i])),
                                              int(self.vox[i]), ctypes.c_void_p(fbase + 8 * i),
                                              single._stream()))
        f = flags.view(n, 2).cpu().tolist()
        return [(2, vmax) if is_int and vmax <= 100 else (fallback, 0) for is_int, vmax in f]

    def stage(self, name: str, match_vol: torch.Tensor, noisy: torch.Tensor, var: torch.Tensor,
              fallback: torch.Tensor, timer=None, pilot: torch.Tensor | None = None,
              ref_thresholds: torch.Tensor | None = None) -> BatchStageOutput:
        L = _lib.lib()
        timer = timer or _NoTimer()
        st = self.stages[name]
        plan = st.plan
        stream = self._single._stream()
        block = (ctypes.c_int32 * 3)(*plan.block3)
        radius = (ctypes.c_int32 * 3)(*plan.radius3)
        R = st.nrefs
        dev = self.device
        ih, idv, nin = st.args()
        sel = torch.empty(R * plan.max_m, dtype=torch.int32, device=dev)
        cnt = torch.empty(R, dtype=torch.int32, device=dev)
        acc = torch.empty(R, dtype=torch.int32, device=dev)
        # one launch per run of consecutive inputs that match in the same mode
        modes = self.input_modes(match_vol, name)
        runs = []
        for i, (mode, tmax) in enumerate(modes):
            if runs and runs[-1][0] == mode:
                runs[-1][1] = max(runs[-1][1], tmax)
                runs[-1][3] = i + 1
            else:
                runs.append([mode, tmax, i, i + 1])
        self.last_runs[name] = [(m, a, b) for m, _, a, b in runs]
        null, null_max = self._single.null_moments() if runs[0][0] >= 7 else (None, 0)
        tok = timer.start(f"{name}.blockmatch")
        for mode, tmax, a, b in runs:
            r0, r1 = int(st.ref_off[a]), int(st.ref_off[b])
            if r1 == r0:
                continue
            off = lambda t, n: ctypes.c_void_p(t.data_ptr() + 4 * n)  # noqa: E731
            _lib.check(L.bmnd_hip_blockmatch_batch(
                _ptr(match_vol), ih, idv, nin, block, radius, off(st.ref_lin, r0),
                off(st.ref_input, r0),
                None if ref_thresholds is None else off(ref_thresholds, r0),
                r1 - r0, plan.threshold, plan.max_m, plan.min_m, 1, 1,
                self._single.count_scale, mode, tmax, _ptr(null), null_max,
                off(sel, r0 * plan.max_m), off(cnt, r0), off(acc, r0), stream))
        timer.stop(tok)
        mats, kai = plan.device_constants(dev)
        k = self._single.k
        kcap = min(k, plan.max_m)
        exact = None
        if kcap > 0:
            exact = torch.empty(R * kcap * plan.P, dtype=torch.float32, device=dev)
            tok = timer.start(f"{name}.exact_std")
            _lib.check(exact_std_entry(name, batch=True)(
                _ptr(var), ih, idv, nin, block, radius, _ptr(sel), _ptr(cnt), _ptr(st.ref_input),
                R, plan.max_m, _ptr(mats), k, _ptr(exact), stream))
            timer.stop(tok)
        filt = torch.empty(R * plan.max_m * plan.P, dtype=torch.float32, device=dev)
        wts = torch.empty(R * plan.max_m, dtype=torch.float32, device=dev)
        wiener = name == "wiener"
        tok = timer.start(f"{name}.filter")
        _lib.check(L.bmnd_hip_filter_groups_batch(
            code_with_transforms(self._single.filter_codes[name], name), _ptr(noisy),
            _ptr((match_vol if pilot is None else pilot) if wiener else None), _ptr(var), ih, idv, nin,
            block, radius, _ptr(sel), _ptr(cnt), _ptr(st.ref_input), R, plan.max_m, _ptr(mats),
            _ptr(exact), plan.shrink, k, self._single.count_scale, _ptr(filt), _ptr(wts), stream))
        timer.stop(tok)
        del exact
        ref_index = torch.empty(st.grid_total, dtype=torch.int32, device=dev)
        _lib.check(L.bmnd_hip_reference_map_batch(_ptr(st.ref_lin), _ptr(st.ref_input), R,
                                                  _ptr(st.inputs_d), _ptr(ref_index),
                                                  st.grid_total, stream))
        est = torch.empty_like(noisy)
        tok = timer.start(f"{name}.aggregate")
        _lib.check(L.bmnd_hip_aggregate_batch(
            _ptr(filt), _ptr(wts), _ptr(sel), _ptr(cnt), _ptr(ref_index), R, plan.max_m, ih, idv,
            nin, block, radius, _ptr(kai), _ptr(fallback), _ptr(est), stream))
        timer.stop(tok)
        del filt, wts, ref_index
        return BatchStageOutput(est, sel.view(R, plan.max_m), cnt, acc, st)

    def ht_thresholds(self, packed: torch.Tensor, timer=None) -> torch.Tensor | None:
        """PoissonPipeline.ht_thresholds for every input, in the flat
        reference order."""
        profile = self._single.profile
        if _ht_policy(profile) is not PoissonHTMatchPolicy.REFERENCE_FINITE_COUNT:
            return None
        t = timer or _NoTimer()
        tok = t.start("ht.thresholds")
        st = self.stages["ht"]
        host = packed.cpu().numpy()
        opts = _threshold_options(profile)
        thr = np.concatenate([
            compute_poisson_reference_thresholds(
                host[self.vox_off[i]:self.vox_off[i + 1]].reshape(self.shapes[i]),
                st.schedules[i], st.plan.block, *opts, distance=self._single.distance)
            for i in range(len(self.shapes))])
        out = torch.from_numpy(thr).to(self.device)
        t.stop(tok)
        return out

    def run_ht(self, packed: torch.Tensor, timer=None):
        """Stage 1 of every input: (HT stage output, stage-1 variance volume,
        None), as PoissonPipeline.run_ht."""
        if (packed.dtype != torch.float32 or not packed.is_contiguous()
                or packed.numel() != self.total_voxels):
            raise ValueError("packed must be a contiguous float32 tensor of every input's voxels")
        t = timer or _NoTimer()
        tok = t.start("ht.variance")
        var1 = self._single.variance(packed)
        t.stop(tok)
        ht = self.stage("ht", packed, packed, var1, packed, timer,
                        ref_thresholds=self.ht_thresholds(packed, t))
        return ht, var1, None

    def run(self, packed: torch.Tensor, timer=None) -> tuple[BatchStageOutput, BatchStageOutput]:
        """packed: float32 device buffer of `total_voxels` (see pack())."""
        t = timer or _NoTimer()
        ht, var1, _ = self.run_ht(packed, timer)
        match2, var2 = self._single._wiener_sources(packed, ht.est, var1, t)
        del var1
        wi = self.stage("wiener", match2, packed, var2, packed, timer, pilot=ht.est)
        return ht, wi

    def graph_modes(self, packed: torch.Tensor) -> tuple[int, int, int]:
        """(stage-1 mode, its table size, stage-2 mode) for capturing the
        batch: the count table when every input holds integer counts <= 100
        (its size the largest count), else the mode stage 1 takes for the
        inputs that do not; stage 2 as PoissonPipeline's graph.  Every mode
        the call-by-call path could take for an input returns the same lists."""
        modes = self.input_modes(packed, "ht")
        if all(m == 2 for m, _ in modes):
            m1, tmax = 2, max(t for _, t in modes)
        else:
            m1, tmax = next((m, t) for m, t in modes if m != 2)
        return m1, tmax, self._single._graph_mode_wiener()


class BatchGraph:
    """Both stages of one or more BatchPipelines (for instance a batch of
    images and a batch of volumes) captured as ONE HIP graph
    (bmnd_hip_pipeline_create_many): a launch reads every batch's packed
    `inputs[j]` and writes `pilots[j]`, `outputs[j]` and each stage's flat
    match lists `lists[j][stage]`, with no allocation or host
    synchronisation between the kernels.  Every output equals the
    BatchPipeline's call-by-call result bit for bit."""

    def __init__(self, pipes):
        self.pipes = list(pipes)
        if not self.pipes:
            raise ValueError("a batch graph needs at least one BatchPipeline")
        devs = {bp.device for bp in self.pipes}
        if len(devs) != 1:
            raise ValueError("every batch of a graph must live on one device")
        self.device = devs.pop()
        for bp in self.pipes:
            if not bp._single.graph_supported():
                raise NotImplementedError("this profile's options run on the call-by-call path")
        dev = self.device
        self.inputs, self.pilots, self.outputs, self.lists = [], [], [], []
        for bp in self.pipes:
            self.inputs.append(torch.empty(bp.total_voxels, dtype=torch.float32, device=dev))
            self.pilots.append(torch.empty_like(self.inputs[-1]))
            self.outputs.append(torch.empty_like(self.inputs[-1]))
            lists = {}
            for name, st in bp.stages.items():
                R = st.nrefs
                lists[name] = (torch.empty(R * st.plan.max_m, dtype=torch.int32, device=dev),
                               torch.empty(R, dtype=torch.int32, device=dev),
                               torch.empty(R, dtype=torch.int32, device=dev))
            self.lists.append(lists)
        self.handle = None
        self.modes = None
        self.k = None

    def _modes(self):
        return [bp.graph_modes(v) for bp, v in zip(self.pipes, self.inputs)]

    def capture(self, modes) -> None:
        self.release()
        descs = (_lib.PipelineDesc * len(self.pipes))()
        for j, bp in enumerate(self.pipes):
            plans = {name: st.plan for name, st in bp.stages.items()}
            refs = {name: st.ref_lin for name, st in bp.stages.items()}
            d = _pipeline_desc(bp._single, plans, refs, self.lists[j], modes[j], self.inputs[j],
                               self.pilots[j], self.outputs[j])
            d.shape[:] = bp.shape3s[0]
            d.ninputs = len(bp.shapes)
            for s, name in enumerate(("ht", "wiener")):
                st = bp.stages[name]
                d.batch[s] = st.inputs_h.ctypes.data
                d.batch_d[s] = st.inputs_d.data_ptr()
                d.ref_input[s] = st.ref_input.data_ptr()
            descs[j] = d
        h = ctypes.c_void_p()
        torch.cuda.synchronize(self.device)
        _lib.check(_lib.lib().bmnd_hip_pipeline_create_many(descs, len(self.pipes),
                                                            ctypes.byref(h)))
        self.handle, self.modes = h, [tuple(m) for m in modes]
        self.k = k_with_order(self.pipes[0]._single.k)

    def launch(self) -> int:
        """One launch on the current stream, then its status (synchronises)."""
        L = _lib.lib()
        stream = self.pipes[0]._single._stream()
        rc = L.bmnd_hip_pipeline_launch(self.handle, stream)
        if rc:
            return rc
        return L.bmnd_hip_pipeline_status(self.handle, stream)

    def run(self, packed=None) -> list[torch.Tensor]:
        """One launch over every batch; `packed` (optional) holds each batch's
        packed input (BatchPipeline.pack), copied into the graph's inputs
        first.  Returns the Wiener estimates (buffers of the graph,
        overwritten by its next launch).  The graph is captured on the first
        call, and again when an input breaks the count-table precondition it
        was captured with (that launch is then repeated with the new graph)."""
        if packed is not None:
            if len(packed) != len(self.pipes):
                raise ValueError("one packed input per batch")
            for dst, src in zip(self.inputs, packed):
                dst.copy_(torch.as_tensor(src).view(-1), non_blocking=False)
        if self.handle is None or self.k != k_with_order(self.pipes[0]._single.k):
            self.capture(self._modes())
        rc = self.launch()
        if rc == _lib_einval():
            self.capture(self._modes())
            rc = self.launch()
        _lib.check(rc)
        return self.outputs

    def release(self) -> None:
        if self.handle is not None:
            torch.cuda.synchronize(self.device)
            _lib.lib().bmnd_hip_pipeline_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass
