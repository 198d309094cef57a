"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

The oracle (``bmnd_oracle.c``) restates the reference denoiser's Poisson
branch on the CPU.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module; the product package
``bmnd_amd`` never does.  What pins each function, and what stays unpinned
(the last bit of ``log``, SSD on non-integer data, the per-voxel sigma-map
values, end-to-end reference outputs), is stated in ``bmnd_oracle.h``.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libbmnd_oracle.so")
_LIB = None

MAX_NDIM = 4
DIST_SSD, DIST_DEVIANCE, DIST_PEARSON, DIST_ANSCOMBE = 0, 1, 2, 3


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


class OracleParams(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int),
        ("shape", ctypes.c_int64 * MAX_NDIM),
        ("ht_block", ctypes.c_int * MAX_NDIM),
        ("ht_step", ctypes.c_int * MAX_NDIM),
        ("ht_radius", ctypes.c_int * MAX_NDIM),
        ("wi_block", ctypes.c_int * MAX_NDIM),
        ("wi_step", ctypes.c_int * MAX_NDIM),
        ("wi_radius", ctypes.c_int * MAX_NDIM),
        ("ht_max", ctypes.c_int),
        ("ht_min", ctypes.c_int),
        ("wi_max", ctypes.c_int),
        ("wi_min", ctypes.c_int),
        ("ht_threshold", ctypes.c_float),
        ("wi_threshold", ctypes.c_float),
        ("ht_lambda", ctypes.c_float),
        ("wi_mu", ctypes.c_float),
        ("ht_beta", ctypes.c_double),
        ("wi_beta", ctypes.c_double),
        ("shift_density", ctypes.c_double),
        ("schedule_density", ctypes.c_int),
        ("k", ctypes.c_int),
        ("count_scale", ctypes.c_double),
        ("variance_floor", ctypes.c_double),
        ("nthreads", ctypes.c_int),
        ("ht_spatial", ctypes.c_void_p),
        ("wi_spatial", ctypes.c_void_p),
        ("wi_match_observation", ctypes.c_int),
        ("wi_var_observation", ctypes.c_int),
        ("ht_thresholds", ctypes.c_void_p),
        ("distance", ctypes.c_int),
        ("ht_standardize", ctypes.c_int),
        ("null_means", ctypes.c_void_p),
        ("null_vars", ctypes.c_void_p),
        ("null_max", ctypes.c_int),
        ("wi_gain_mode", ctypes.c_int),
        ("ht_soft", ctypes.c_int),
        ("ht_wmodel", ctypes.c_int),
        ("wi_wmodel", ctypes.c_int),
        ("ht_windowed", ctypes.c_int),
        ("wi_windowed", ctypes.c_int),
        ("ht_group_scope", ctypes.c_int),
        ("wi_group_scope", ctypes.c_int),
        ("ht_smd", ctypes.c_void_p),
        ("wi_smd", ctypes.c_void_p),
        ("ht_gmd", ctypes.c_void_p),
        ("wi_gmd", ctypes.c_void_p),
        ("mass_mode", ctypes.c_int),
        ("ht_mdir", ctypes.c_void_p),
        ("wi_mdir", ctypes.c_void_p),
        ("ht_sdir", ctypes.c_void_p),
        ("wi_sdir", ctypes.c_void_p),
        ("gaussian", ctypes.c_int),
        ("ht_match_d", ctypes.c_double),
        ("wi_match_d", ctypes.c_double),
        ("ht_cov", ctypes.c_void_p),
        ("wi_cov", ctypes.c_void_p),
        ("cov_domain", ctypes.c_int * MAX_NDIM),
        ("ht_noise_ssd", ctypes.c_void_p),
        ("ht_gamma", ctypes.c_float),
        ("global_sigma", ctypes.c_double),
        ("poisson_ssd", ctypes.c_int),
    ]


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        build()
        _LIB = ctypes.CDLL(_LIB_PATH)
        _declare(_LIB)
    return _LIB


def _p(a: np.ndarray | None) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if a is None else a.ctypes.data)


def _declare(L: ctypes.CDLL) -> None:
    vp, i32, i64, f32, f64 = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                              ctypes.c_float, ctypes.c_double)
    L.oracle_kaiser.argtypes = [i32, f64, vp]
    L.oracle_kaiser_nd.argtypes = [i32, vp, f64, vp]
    L.oracle_haar_matrix.argtypes = [i32, vp]
    L.oracle_haar_matrix.restype = i32
    L.oracle_dct_matrix.argtypes = [i32, vp, vp]
    L.oracle_bior15_matrix.argtypes = [i32, vp, vp]
    L.oracle_bior15_matrix.restype = i32
    L.oracle_schedule_generated.argtypes = [i32, vp, vp, vp, f64, i32, vp, i64]
    L.oracle_schedule_generated.restype = i64
    L.oracle_canonical_key.argtypes = [f32]
    L.oracle_canonical_key.restype = f32
    L.oracle_patch_distance.argtypes = [vp, vp, i32, i32, f64]
    L.oracle_patch_distance.restype = f32
    L.oracle_blockmatch.argtypes = [vp, vp, i32, vp, vp, vp, vp, i64, vp, vp,
                                    i32, i32, i32, i32, i32, f64, i32, vp, vp, vp]
    L.oracle_standardized_distance.argtypes = [vp, vp, i32, i32, f64, vp, vp, i32]
    L.oracle_standardized_distance.restype = f32
    L.oracle_blockmatch_standardized.argtypes = [vp, vp, i32, vp, vp, vp, vp, i64, vp, vp,
                                                 i32, i32, i32, i32, i32, f64, vp, vp, i32, i32,
                                                 vp, vp, vp]
    L.oracle_blockmatch_ssd.argtypes = [vp, i32, vp, vp, vp, vp, i64, vp, vp, i32, i32, i32,
                                        vp, f32, i32, vp, vp, vp]
    L.oracle_blockmatch_ssd_poisson.argtypes = [vp, i32, vp, vp, vp, vp, i64, vp, vp, i32, i32,
                                                i32, f32, f64, i32, vp, vp, vp]
    L.oracle_gaussian_group_variance.argtypes = [vp, i32, i64, vp, vp, i32, i32, vp]
    L.oracle_ht_gaussian.argtypes = [vp, vp, i64, f32, vp]
    L.oracle_wiener_gaussian.argtypes = [vp, vp, vp, i64, f32, vp, vp]
    L.oracle_apply_transform.argtypes = [vp, i32, vp, vp]
    L.oracle_exact_variance.argtypes = [vp, i32, vp, vp, i32, vp, vp, vp, i32, vp]
    L.oracle_exact_variance.restype = i32
    L.oracle_noise_std.argtypes = [vp, i32, vp, vp, i32, vp, vp, vp, i32, vp]
    L.oracle_noise_std.restype = i32
    L.oracle_limit_planes.argtypes = [vp, vp, i32, i32, i32, vp]
    L.oracle_patch_weights.argtypes = [vp, vp, i32, i32, vp, f64, vp]
    L.oracle_ht_shrink.argtypes = [vp, vp, i64, f32, f64, vp]
    L.oracle_ht_soft.argtypes = [vp, vp, i64, f32, vp]
    L.oracle_np_sum_f64.argtypes = [vp, i64]
    L.oracle_np_sum_f64.restype = f64
    L.oracle_weight_risks.argtypes = [vp, vp, vp, i32, i32, vp, vp, vp, i32, i32, i32, vp]
    L.oracle_weight_risks.restype = i32
    L.oracle_weights_ex.argtypes = [vp, vp, vp, i32, i32, vp, vp, vp, i32, i32, i32, f64, vp]
    L.oracle_exact_group_risk.argtypes = [vp, i32, vp, vp, i32, vp, vp, vp, vp, vp, vp, i32, vp,
                                          vp, vp]
    L.oracle_exact_group_risk.restype = f64
    L.oracle_mass_correct.argtypes = [i32, vp, vp, vp, vp, vp, vp, f64]
    L.oracle_wiener_shrink.argtypes = [vp, vp, vp, i64, f32, vp]
    L.oracle_wiener_shrink_mode.argtypes = [vp, vp, vp, i64, f32, i32, vp]
    L.oracle_scatter_add.argtypes = [vp, vp, i32, vp, vp, vp, vp, vp, vp, i32]
    L.oracle_bmnd_poisson.argtypes = [vp, ctypes.POINTER(OracleParams), vp, vp,
                                      vp, vp, i64, vp, vp, i64]
    L.oracle_bmnd_poisson.restype = i32
    L.oracle_bmnd_poisson_sigma.argtypes = [vp, ctypes.POINTER(OracleParams), vp, vp,
                                            vp, vp, i64, vp, vp, i64, vp, vp, vp]
    L.oracle_bmnd_poisson_sigma.restype = i32
    L.oracle_np_sum_f32.argtypes = [vp, i64]
    L.oracle_np_sum_f32.restype = f32
    L.oracle_np_mean_f32.argtypes = [vp, i64]
    L.oracle_np_mean_f32.restype = f32
    L.oracle_poisson_global_sigma.argtypes = [vp, i64, f64]
    L.oracle_poisson_global_sigma.restype = f64
    L.oracle_set_exact_order.argtypes = [i32]
    L.oracle_numba_argsort.argtypes = [vp, i64, vp]


# ---------------------------------------------------------------------------
# thin numpy wrappers
# ---------------------------------------------------------------------------

def kaiser(m: int, beta: float) -> np.ndarray:
    out = np.empty(m, dtype=np.float64)
    lib().oracle_kaiser(m, beta, _p(out))
    return out


def kaiser_nd(block: tuple[int, ...], beta: float) -> np.ndarray:
    b = np.asarray(block, dtype=np.int32)
    out = np.empty(int(np.prod(block)), dtype=np.float32)
    lib().oracle_kaiser_nd(len(block), _p(b), beta, _p(out))
    return out.reshape(block)


def haar(n: int) -> np.ndarray:
    out = np.empty((n, n), dtype=np.float32)
    if lib().oracle_haar_matrix(n, _p(out)) != 0:
        raise ValueError(f"Haar size must be a power of two, got {n}")
    return out


def dct(n: int) -> tuple[np.ndarray, np.ndarray]:
    f = np.empty((n, n), dtype=np.float32)
    i = np.empty((n, n), dtype=np.float32)
    lib().oracle_dct_matrix(n, _p(f), _p(i))
    return f, i


def bior15(n: int) -> tuple[np.ndarray, np.ndarray]:
    f = np.empty((n, n), dtype=np.float32)
    i = np.empty((n, n), dtype=np.float32)
    if lib().oracle_bior15_matrix(n, _p(f), _p(i)) != 0:
        raise ValueError(f"no official bior1.5 matrix of size {n}")
    return f, i


def schedule(counts, block, step, shift_density=2.0, schedule_density=2) -> np.ndarray:
    ndim = len(counts)
    c = np.asarray(counts, dtype=np.int64)
    b = np.asarray(block, dtype=np.int32)
    s = np.asarray(step, dtype=np.int32)
    n = lib().oracle_schedule_generated(ndim, _p(c), _p(b), _p(s), float(shift_density),
                                        int(schedule_density), None, 0)
    out = np.empty((n, ndim), dtype=np.int64)
    lib().oracle_schedule_generated(ndim, _p(c), _p(b), _p(s), float(shift_density),
                                    int(schedule_density), _p(out), n)
    return out


def numba_argsort(a) -> np.ndarray:
    """argsort in numba's default (quicksort) order, from the C oracle."""
    arr = np.ascontiguousarray(a, dtype=np.int64)
    out = np.empty(arr.size, dtype=np.int32)
    lib().oracle_numba_argsort(_p(arr), arr.size, _p(out))
    return out


def set_exact_order(quicksort: bool) -> None:
    """Union-voxel entry order of the exact variance: stable (default) or
    numba's quicksort argsort order as the reference leaves it."""
    lib().oracle_set_exact_order(int(bool(quicksort)))


def canonical_key(d: float) -> float:
    return float(lib().oracle_canonical_key(float(np.float32(d))))


def patch_distance(ref, cand, code=DIST_DEVIANCE, count_scale=1.0) -> np.float32:
    r = np.ascontiguousarray(ref, dtype=np.float32)
    c = np.ascontiguousarray(cand, dtype=np.float32)
    return np.float32(lib().oracle_patch_distance(_p(r), _p(c), r.size, code, count_scale))


@dataclass
class MatchResult:
    selected: np.ndarray  # (nrefs, max) int64, -1 padded
    counts: np.ndarray  # (nrefs,) int32
    accepted: np.ndarray  # (nrefs,) int64

    def lists(self) -> list[list[int]]:
        return [self.selected[i, : self.counts[i]].tolist() for i in range(len(self.counts))]


def blockmatch(refs, counts, radius, max_matches, min_matches, threshold, *,
               patches_flat=None, volume=None, block=None, pow2=True, stabilize=True,
               code=DIST_DEVIANCE, count_scale=1.0, nthreads=0,
               null_moments=None) -> MatchResult:
    """Stage matching of every reference (blockmatching.py:1045-1205); with
    null_moments = (float32 means, float32 variances) the candidate-
    standardized distance."""
    refs = np.ascontiguousarray(np.asarray(refs, dtype=np.int64).reshape(-1, len(counts)))
    nrefs, ndim = refs.shape
    cts = np.asarray(counts, dtype=np.int64)
    rad = np.asarray(radius, dtype=np.int32)
    thr = np.asarray(threshold, dtype=np.float32)
    if thr.ndim == 0:
        thr = np.full(nrefs, thr, dtype=np.float32)
    thr = np.ascontiguousarray(thr)
    if patches_flat is not None:
        pf = np.ascontiguousarray(patches_flat, dtype=np.float32)
        vol, vshape = None, None
        blk = np.ones(ndim, dtype=np.int32)
        blk[0] = pf.shape[1]
    else:
        pf = None
        vol = np.ascontiguousarray(volume, dtype=np.float32)
        vshape = np.asarray(vol.shape, dtype=np.int64)
        blk = np.asarray(block, dtype=np.int32)
    sel = np.empty((nrefs, max_matches), dtype=np.int64)
    cnt = np.empty(nrefs, dtype=np.int32)
    acc = np.empty(nrefs, dtype=np.int64)
    if null_moments is None:
        lib().oracle_blockmatch(_p(pf), _p(vol), ndim, _p(vshape), _p(cts), _p(blk), _p(refs),
                                nrefs, _p(rad), _p(thr), max_matches, min_matches, int(pow2),
                                int(stabilize), code, float(count_scale), nthreads, _p(sel),
                                _p(cnt), _p(acc))
    else:
        nm = np.ascontiguousarray(null_moments[0], dtype=np.float32)
        nv = np.ascontiguousarray(null_moments[1], dtype=np.float32)
        lib().oracle_blockmatch_standardized(
            _p(pf), _p(vol), ndim, _p(vshape), _p(cts), _p(blk), _p(refs), nrefs, _p(rad),
            _p(thr), max_matches, min_matches, int(pow2), int(stabilize), code,
            float(count_scale), _p(nm), _p(nv), len(nm) - 1, nthreads, _p(sel), _p(cnt), _p(acc))
    return MatchResult(sel, cnt, acc)


def apply_transform(data: np.ndarray, mats: list) -> np.ndarray:
    d = np.ascontiguousarray(data, dtype=np.float32).copy()
    dims = np.asarray(d.shape, dtype=np.int32)
    keep = [None if m is None else np.ascontiguousarray(m, dtype=np.float32) for m in mats]
    ptrs = (ctypes.c_void_p * len(keep))(*[None if m is None else m.ctypes.data for m in keep])
    lib().oracle_apply_transform(_p(d), d.ndim, _p(dims), ptrs)
    return d


def _spatial_ptrs(spatial):
    keep = [None if m is None else np.ascontiguousarray(m, dtype=np.float32) for m in spatial]
    ptrs = (ctypes.c_void_p * max(len(keep), 1))(
        *[None if m is None else m.ctypes.data for m in keep])
    return keep, ptrs


def exact_variance(var_volume, block, abs_pos, group_mat, spatial, planes) -> np.ndarray:
    vv = np.ascontiguousarray(var_volume, dtype=np.float32)
    ndim = vv.ndim
    shp = np.asarray(vv.shape, dtype=np.int64)
    blk = np.asarray(block, dtype=np.int32)
    ap = np.ascontiguousarray(abs_pos, dtype=np.int64)
    g = np.ascontiguousarray(group_mat, dtype=np.float32)
    keep, ptrs = _spatial_ptrs(spatial)
    P = int(np.prod(block))
    out = np.empty(planes * P, dtype=np.float32)
    lib().oracle_exact_variance(_p(vv), ndim, _p(shp), _p(blk), ap.shape[0], _p(ap), _p(g),
                                ptrs, planes, _p(out))
    return out.reshape((planes,) + tuple(block))


def noise_std(var_volume, block, abs_pos, group_mat, spatial, k) -> np.ndarray:
    vv = np.ascontiguousarray(var_volume, dtype=np.float32)
    ndim = vv.ndim
    shp = np.asarray(vv.shape, dtype=np.int64)
    blk = np.asarray(block, dtype=np.int32)
    ap = np.ascontiguousarray(abs_pos, dtype=np.int64)
    n = ap.shape[0]
    g = np.ascontiguousarray(group_mat, dtype=np.float32)
    keep, ptrs = _spatial_ptrs(spatial)
    P = int(np.prod(block))
    out = np.empty(n * P, dtype=np.float32)
    lib().oracle_noise_std(_p(vv), ndim, _p(shp), _p(blk), n, _p(ap), _p(g), ptrs, k, _p(out))
    return out.reshape((n,) + tuple(block))


def limit_planes(exact_std, approx_std, k) -> np.ndarray:
    a = np.ascontiguousarray(approx_std, dtype=np.float32)
    n = a.shape[0]
    p = int(a.size // n)
    e = np.zeros_like(a)
    ex = np.ascontiguousarray(exact_std, dtype=np.float32)
    e[: ex.shape[0]] = ex
    out = np.empty_like(a)
    lib().oracle_limit_planes(_p(e), _p(a), n, p, k, _p(out))
    return out


def patch_weights(atten, sigma, group_inv, eps=1e-10) -> np.ndarray:
    a = np.ascontiguousarray(atten, dtype=np.float32)
    s = np.ascontiguousarray(sigma, dtype=np.float32)
    n = a.shape[0]
    p = int(a.size // n)
    gi = np.ascontiguousarray(group_inv, dtype=np.float32)
    out = np.empty(n, dtype=np.float32)
    lib().oracle_patch_weights(_p(a), _p(s), n, p, _p(gi), eps, _p(out))
    return out


def ht_shrink(coef, sigma, lam=3.0, count_scale=1.0):
    c = np.ascontiguousarray(coef, dtype=np.float32).copy()
    s = np.ascontiguousarray(sigma, dtype=np.float32)
    att = np.empty_like(c)
    lib().oracle_ht_shrink(_p(c), _p(s), c.size, lam, count_scale, _p(att))
    return c, att


def ht_soft(coef, sigma, lam=3.0):
    """Soft threshold: (coefficients, attenuation)."""
    c = np.ascontiguousarray(coef, dtype=np.float32).copy()
    s = np.ascontiguousarray(sigma, dtype=np.float32)
    att = np.empty_like(c)
    lib().oracle_ht_soft(_p(c), _p(s), c.size, lam, _p(att))
    return c, att


def np_sum_f64(a) -> float:
    x = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    return float(lib().oracle_np_sum_f64(_p(x), x.size))


WEIGHT_MODELS = {"variance": 0, "classic": 1, "risk": 2}


def weight_risks(atten, sigma, group_inv, *, model="variance", signal=None, smd=None, gmd=None,
                 windowed=False, group_scope=False) -> np.ndarray:
    """The diagonal-filter aggregation risks (n for patch scope, 1 for group
    scope) of an (n, P) group; gmd is the group metric diagonal of this n."""
    a = np.ascontiguousarray(atten, dtype=np.float32)
    n = a.shape[0]
    p = int(a.size // n)
    s = np.ascontiguousarray(sigma, dtype=np.float32)
    sp = np.ascontiguousarray(signal if signal is not None else np.zeros_like(a), dtype=np.float32)
    gi = np.ascontiguousarray(group_inv, dtype=np.float32)
    sm = None if smd is None else np.ascontiguousarray(smd, dtype=np.float64)
    gm = None if gmd is None else np.ascontiguousarray(gmd, dtype=np.float64)
    out = np.empty(n, dtype=np.float64)
    k = lib().oracle_weight_risks(_p(a), _p(s), _p(sp), n, p, _p(gi), _p(sm), _p(gm),
                                  WEIGHT_MODELS[model], int(windowed), int(group_scope), _p(out))
    return out[:k]


def exact_group_risk(var_volume, block, abs_pos, group_fwd, group_inv, spatial_fwd, spatial_inv,
                     window, planes, gain, sigma, spow=None) -> float:
    """The exact covariance-aware group risk (wiener.py:638-844) of one group:
    exact over the first `planes` planes, diagonal over the rest, plus the
    bias when `spow` (the signal power) is given.  gain, sigma, spow: (n, P)."""
    vv = np.ascontiguousarray(var_volume, dtype=np.float32)
    ndim = vv.ndim
    shp = np.asarray(vv.shape, dtype=np.int64)
    blk = np.asarray(block, dtype=np.int32)
    ap = np.ascontiguousarray(abs_pos, dtype=np.int64)
    n = ap.shape[0]
    gf = np.ascontiguousarray(group_fwd, dtype=np.float32)
    gi = np.ascontiguousarray(group_inv, dtype=np.float32)
    keep_f, fptrs = _spatial_ptrs(spatial_fwd)
    keep_i, iptrs = _spatial_ptrs(spatial_inv)
    w = np.ascontiguousarray(window, dtype=np.float32).reshape(-1)
    g = np.ascontiguousarray(gain, dtype=np.float32).reshape(n, -1)
    s = np.ascontiguousarray(sigma, dtype=np.float32).reshape(n, -1)
    sp = None if spow is None else np.ascontiguousarray(spow, dtype=np.float32).reshape(n, -1)
    return float(lib().oracle_exact_group_risk(_p(vv), ndim, _p(shp), _p(blk), n, _p(ap), _p(gf),
                                               _p(gi), fptrs, iptrs, _p(w), int(planes), _p(g),
                                               _p(s), _p(sp)))


GAIN_MODES = {"variance-scaled": 0, "classic": 1, "noise-floor": 2}


def wiener_shrink(coef, ref_coef, sigma, mu=0.4, mode="variance-scaled"):
    c = np.ascontiguousarray(coef, dtype=np.float32).copy()
    r = np.ascontiguousarray(ref_coef, dtype=np.float32)
    s = np.ascontiguousarray(sigma, dtype=np.float32)
    att = np.empty_like(c)
    lib().oracle_wiener_shrink_mode(_p(c), _p(r), _p(s), c.size, mu, GAIN_MODES[mode], _p(att))
    return c, att


def scatter_add(num, den, block, kaiser, filt, weights, lins):
    """In-place sequential scatter of one group into num/den."""
    shp = np.asarray(num.shape, dtype=np.int64)
    blk = np.asarray(block, dtype=np.int32)
    k = np.ascontiguousarray(kaiser, dtype=np.float32)
    f = np.ascontiguousarray(filt, dtype=np.float32)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    ln = np.ascontiguousarray(lins, dtype=np.int64)
    assert num.dtype == np.float32 and num.flags.c_contiguous and den.flags.c_contiguous
    lib().oracle_scatter_add(_p(num), _p(den), num.ndim, _p(shp), _p(blk), _p(k), _p(f), _p(w),
                             _p(ln), ln.size)


# the official bior1.5 analysis tables of transforms.py:711-806 (data)
_B8 = (0.343550200747110, -0.225454819240296, -0.461645582253923,
       0.569359398342840, 0.402347308162280, -0.083506045090280, 0.707106781186550)


def reference_spatial_matrices(kind: str, block) -> np.ndarray:
    """Analysis then synthesis matrices of each axis, flattened, built the way
    the reference builds them with numpy: the DCT-II from float64 cosines and
    the bior1.5 synthesis as numpy.linalg.inv of the float32 table
    (transforms.py:359-386, 699-811)."""
    fwd, inv = [], []
    for b in block:
        if kind == "dct":
            m = np.zeros((b, b), np.float32)
            for k in range(b):
                norm = np.sqrt(1 / b) if k == 0 else np.sqrt(2 / b)
                for n in range(b):
                    m[k, n] = norm * np.cos(np.pi * (n + 0.5) * k / b)
            fwd.append(m)
            inv.append(np.ascontiguousarray(m.T))
        else:
            f, _ = bior15(b)  # the C table (checked bitwise against the literal)
            fwd.append(f)
            inv.append(np.asarray(np.linalg.inv(f), dtype=np.float32))
    return np.concatenate([m.reshape(-1) for m in fwd + inv]).astype(np.float32)


def default_params(ndim: int, shape, nthreads: int = 0) -> OracleParams:
    """Default BM3DProfile / BM4DProfile values with noise_model=POISSON
    (profiles.py:225-352)."""
    p = OracleParams()
    p.ndim = ndim
    for d in range(ndim):
        p.shape[d] = int(shape[d])
    if ndim == 2:
        hb, wb, rad, wthr = (8, 8), (8, 8), (19, 19), 0.3937
    elif ndim == 3:
        hb, wb, rad, wthr = (4, 4, 4), (5, 5, 5), (7, 7, 7), 0.7689
    else:
        raise ValueError("default Poisson profiles exist for 2-D and 3-D inputs")
    for d in range(ndim):
        p.ht_block[d], p.ht_step[d], p.ht_radius[d] = hb[d], 3, rad[d]
        p.wi_block[d], p.wi_step[d], p.wi_radius[d] = wb[d], 3, rad[d]
    p.ht_max, p.ht_min, p.wi_max, p.wi_min = 16, 2, 32, 2
    p.ht_threshold, p.wi_threshold = 2.9527, wthr
    p.ht_lambda, p.wi_mu = 3.0, 0.4
    p.ht_beta = p.wi_beta = 2.0
    p.shift_density, p.schedule_density = 2.0, 2
    p.k = 4
    p.count_scale = 1.0
    p.variance_floor = 1e-10
    p.distance = DIST_DEVIANCE
    p.nthreads = nthreads
    p._keep = (reference_spatial_matrices("bior", hb), reference_spatial_matrices("dct", wb))
    p.ht_spatial = p._keep[0].ctypes.data
    p.wi_spatial = p._keep[1].ctypes.data
    return p


# ---- the FP64 build (libbmnd_oracle_f64.so): the same algorithm with every
# float32 quantity in float64, for the error of the float32 path
_LIB64_PATH = os.path.join(_HERE, "libbmnd_oracle_f64.so")
_LIB64 = None

OracleParams64 = type("OracleParams64", (ctypes.Structure,), {
    "_fields_": [(n, ctypes.c_double if t is ctypes.c_float else t)
                 for n, t in OracleParams._fields_]})


def lib64() -> ctypes.CDLL:
    global _LIB64
    if _LIB64 is None:
        if not os.path.exists(_LIB64_PATH):
            build(force=True)
        L = ctypes.CDLL(_LIB64_PATH)
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        L.oracle_bmnd_poisson.argtypes = [vp, ctypes.POINTER(OracleParams64), vp, vp, vp, vp, i64,
                                          vp, vp, i64]
        L.oracle_bmnd_poisson.restype = i32
        L.oracle_dct_matrix.argtypes = [ctypes.c_int, vp, vp]
        L.oracle_dct_matrix.restype = None
        L.oracle_bior15_matrix.argtypes = [ctypes.c_int, vp, vp]
        L.oracle_bior15_matrix.restype = ctypes.c_int
        _LIB64 = L
    return _LIB64


def _spatial_matrices_f64(kind: str, block) -> np.ndarray:
    """Analysis then synthesis matrices of each axis in float64: the DCT-II
    from float64 cosines, bior1.5 from its decimal table with a float64
    inverse (the FP64 build's own constructions)."""
    L = lib64()
    fwd, inv = [], []
    for b in block:
        f = np.empty((b, b), np.float64)
        i = np.empty((b, b), np.float64)
        if kind == "dct":
            L.oracle_dct_matrix(b, f.ctypes.data, i.ctypes.data)
        elif L.oracle_bior15_matrix(b, f.ctypes.data, i.ctypes.data) != 0:
            raise ValueError(f"no bior1.5 matrix of size {b}")
        fwd.append(f)
        inv.append(i)
    return np.concatenate([m.reshape(-1) for m in fwd + inv])


def default_params_f64(ndim: int, shape, nthreads: int = 0) -> OracleParams64:
    """default_params() for the FP64 build."""
    p32 = default_params(ndim, shape, nthreads)
    p = OracleParams64()
    for name, _ in OracleParams._fields_:
        if name not in ("ht_spatial", "wi_spatial", "ht_thresholds", "null_means", "null_vars"):
            setattr(p, name, getattr(p32, name))
    hb = tuple(p32.ht_block[d] for d in range(ndim))
    wb = tuple(p32.wi_block[d] for d in range(ndim))
    p._keep = (_spatial_matrices_f64("bior", hb), _spatial_matrices_f64("dct", wb))
    p.ht_spatial = p._keep[0].ctypes.data
    p.wi_spatial = p._keep[1].ctypes.data
    return p


def bmnd_poisson_f64(vol: np.ndarray, params: OracleParams64 | None = None, *,
                     nthreads: int = 0):
    """The two-stage pipeline computed in float64 throughout: (out, pilot)
    as float64 arrays."""
    v = np.ascontiguousarray(vol, dtype=np.float64)
    if params is None:
        params = default_params_f64(v.ndim, v.shape, nthreads)
    out = np.empty_like(v)
    pilot = np.empty_like(v)
    rc = lib64().oracle_bmnd_poisson(_p(v), ctypes.byref(params), _p(out), _p(pilot), None, None,
                                     0, None, None, 0)
    if rc != 0:
        raise RuntimeError(f"oracle_bmnd_poisson (float64) failed with code {rc}")
    return out, pilot


def np_sum_f32(a: np.ndarray) -> np.float32:
    """numpy's float32 summation order, restated (np.sum of a contiguous array)."""
    v = np.ascontiguousarray(a, dtype=np.float32).ravel()
    return np.float32(lib().oracle_np_sum_f32(_p(v), v.size))


def np_mean_f32(a: np.ndarray) -> np.float32:
    v = np.ascontiguousarray(a, dtype=np.float32).ravel()
    return np.float32(lib().oracle_np_mean_f32(_p(v), v.size))


def poisson_global_sigma(var_volume: np.ndarray, floor: float = 1e-10) -> float:
    v = np.ascontiguousarray(var_volume, dtype=np.float32).ravel()
    return float(lib().oracle_poisson_global_sigma(_p(v), v.size, floor))


def bmnd_poisson_sigma_maps(vol: np.ndarray, params: OracleParams | None = None, *,
                            nthreads: int = 0):
    """The oracle's pipeline with the reference's ``return_sigma_map`` outputs:
    ``(out, {"global": g, "ht": map, "wiener": map})`` (maps of zeros and no
    "global" entry for an all-zero input, as the reference returns)."""
    v = np.ascontiguousarray(vol, dtype=np.float32)
    if params is None:
        params = default_params(v.ndim, v.shape, nthreads)
    out = np.empty_like(v)
    hts, wis = np.empty_like(v), np.empty_like(v)
    g = ctypes.c_double(0.0)
    rc = lib().oracle_bmnd_poisson_sigma(_p(v), ctypes.byref(params), _p(out), None,
                                         None, None, 0, None, None, 0, _p(hts), _p(wis),
                                         ctypes.byref(g))
    if rc != 0:
        raise RuntimeError(f"oracle_bmnd_poisson_sigma failed with code {rc}")
    if not np.any(v != 0.0):
        return out, {"ht": hts, "wiener": wis}
    return out, {"global": g.value, "ht": hts, "wiener": wis}


def bmnd_poisson(vol: np.ndarray, params: OracleParams | None = None, *, nthreads: int = 0,
                 capture_matches: bool = False):
    """Run the oracle's two-stage Poisson pipeline.

    Returns ``(out, pilot)`` or, with ``capture_matches``, ``(out, pilot,
    (ht_sel, ht_cnt), (wi_sel, wi_cnt))`` where the selections are (nrefs, max)
    arrays of linear patch indices padded with -1."""
    v = np.ascontiguousarray(vol, dtype=np.float32)
    if params is None:
        params = default_params(v.ndim, v.shape, nthreads)
    out = np.empty_like(v)
    pilot = np.empty_like(v)
    ht = wi = (None, None)
    args = [None, None, 0, None, None, 0]
    if capture_matches:
        nd = v.ndim
        cnt_ht = tuple(int(v.shape[d]) - params.ht_block[d] + 1 for d in range(nd))
        cnt_wi = tuple(int(v.shape[d]) - params.wi_block[d] + 1 for d in range(nd))
        n_ht = len(schedule(cnt_ht, tuple(params.ht_block[:nd]), tuple(params.ht_step[:nd]),
                            params.shift_density, params.schedule_density))
        n_wi = len(schedule(cnt_wi, tuple(params.wi_block[:nd]), tuple(params.wi_step[:nd]),
                            params.shift_density, params.schedule_density))
        ht = (np.empty((n_ht, params.ht_max), np.int64), np.empty(n_ht, np.int32))
        wi = (np.empty((n_wi, params.wi_max), np.int64), np.empty(n_wi, np.int32))
        args = [ht[0], ht[1], n_ht, wi[0], wi[1], n_wi]
    rc = lib().oracle_bmnd_poisson(_p(v), ctypes.byref(params), _p(out), _p(pilot),
                                   _p(args[0]), _p(args[1]), args[2], _p(args[3]),
                                   _p(args[4]), args[5])
    if rc != 0:
        raise RuntimeError(f"oracle_bmnd_poisson failed with code {rc}")
    if capture_matches:
        return out, pilot, ht, wi
    return out, pilot


# ---- per-reference stage-1 thresholds ("reference_finite_count" policy) ----
# Pure-Python loops in the reference's literal order, for small inputs.

def poisson_null_moments(max_count: int = 512, code: int = DIST_DEVIANCE) -> tuple[list, list]:
    """Mean and variance (float64) of one distance term when (r, c) is a
    Binomial(N, 1/2) split of a pooled count N, for N = 0..max_count; N = 0
    stays 0 (blockmatching.py:29-87).  Deviance term 2[r log(2r/N) +
    c log(2c/N)]; Anscombe term (2 sqrt(c + 3/8) - 2 sqrt(r + 3/8))^2 / 2;
    Pearson: mean 1 and variance 2 - 2/N in closed form."""
    import math
    means, variances = [0.0] * (max_count + 1), [0.0] * (max_count + 1)
    if code == DIST_PEARSON:
        for n in range(1, max_count + 1):
            means[n], variances[n] = 1.0, 2.0 - 2.0 / float(n)
        return means, variances
    for n in range(1, max_count + 1):
        e = s2 = 0.0
        half_n = 2.0 ** (-n)
        for r in range(n + 1):
            c = n - r
            p = math.comb(n, r) * half_n
            if code == DIST_DEVIANCE:
                t = 0.0
                if r > 0:
                    t += r * math.log(2.0 * r / n)
                if c > 0:
                    t += c * math.log(2.0 * c / n)
                t *= 2.0
            else:
                t = 0.5 * (2.0 * math.sqrt(c + 0.375) - 2.0 * math.sqrt(r + 0.375)) ** 2
            e += p * t
            s2 += p * t * t
        means[n] = e
        variances[n] = max(s2 - e * e, 0.0)
    return means, variances


def poisson_reference_thresholds(vol: np.ndarray, params, ht_match_threshold: float,
                                 structure_beta: float, intensity_power: float,
                                 max_count: int = 512) -> np.ndarray:
    """float32 stage-1 threshold of every stage-1 reference, in schedule
    order (blockmatching.py:90-148): the mean of the reference patch's null
    means, plus ht_match_threshold times sqrt(sum of null variances) / P,
    plus structure_beta times the mean count to intensity_power; pooled
    counts above max_count take the chi-square(1) moments 1 and 2."""
    import math
    v = np.ascontiguousarray(vol, dtype=np.float32)
    nd = v.ndim
    block = tuple(params.ht_block[:nd])
    counts = tuple(int(v.shape[d]) - block[d] + 1 for d in range(nd))
    refs = schedule(counts, block, tuple(params.ht_step[:nd]), params.shift_density,
                    params.schedule_density)
    means, variances = poisson_null_moments(max_count, params.distance or DIST_DEVIANCE)
    P = int(np.prod(block))
    out = np.empty(len(refs), dtype=np.float32)
    for i, ref in enumerate(refs):
        patch = v[tuple(slice(int(ref[d]), int(ref[d]) + block[d]) for d in range(nd))]
        cnt = np.clip(patch.reshape(-1).astype(np.float64) / float(params.count_scale), 0.0, None)
        nm = np.empty(P)
        nv = np.empty(P)
        for j in range(P):
            pooled = np.rint(2.0 * cnt[j])
            if pooled > max_count:
                nm[j], nv[j] = 1.0, 2.0
            else:
                nm[j], nv[j] = means[int(pooled)], variances[int(pooled)]
        thr = (float(np.mean(nm)) + float(ht_match_threshold)
               * math.sqrt(max(float(np.sum(nv)) / (P * P), 0.0))
               + float(structure_beta) * float(np.mean(cnt)) ** float(intensity_power))
        out[i] = np.float32(thr)
    return out


def group_inverse(n: int) -> np.ndarray:
    """Synthesis matrix of the member axis: the transpose of the Haar
    matrix for powers of two, of the DCT-II otherwise (transforms.py:411-440)."""
    f = haar(n) if n & (n - 1) == 0 else dct(n)[0]
    return np.ascontiguousarray(f.T)


def _stage_synthesis(params, stage):
    nd = params.ndim
    block = tuple(params.ht_block[:nd] if stage == "ht" else params.wi_block[:nd])
    mats = params._keep[0 if stage == "ht" else 1]
    o = sum(b * b for b in block)
    inv = []
    for b in block:
        inv.append(np.asarray(mats[o:o + b * b], np.float32).reshape(b, b))
        o += b * b
    beta = params.ht_beta if stage == "ht" else params.wi_beta
    return block, inv, kaiser_nd(block, beta)


def synthesis_metric_tables(params, stage: str):
    """A stage's spatial metric diagonal [P] and its group metric diagonals
    (n entries for n = 1..max, concatenated), as
    aggregationweights.py:79-120 builds them: the Kronecker product of the
    spatial synthesis matrices S in float64, diag(S^T diag(w^2) S) with the
    float32 Kaiser window w, and diag(Ginv^T Ginv)."""
    _, inv, window = _stage_synthesis(params, stage)
    synth = np.array([[1.0]], dtype=np.float64)
    for m in inv:
        synth = np.kron(synth, m.astype(np.float64))
    w = window.astype(np.float64).reshape(-1)
    smd = np.diag(synth.T @ ((w ** 2)[:, None] * synth)).copy()
    max_m = params.ht_max if stage == "ht" else params.wi_max
    gmd = []
    for n in range(1, max_m + 1):
        gi = group_inverse(n).astype(np.float64)
        gmd.append(np.diag(gi.T @ gi).copy())
    return smd, np.concatenate(gmd)


def set_weight_options(params, stage: str, model: str = "variance", windowed: bool = False,
                       group_scope: bool = False) -> None:
    """A stage's aggregation weight model, domain and scope
    (aggregationweights.py:123-259), with the metric tables it reads."""
    pre = "ht_" if stage == "ht" else "wi_"
    setattr(params, pre + "wmodel", WEIGHT_MODELS[model])
    setattr(params, pre + "windowed", int(windowed))
    setattr(params, pre + "group_scope", int(group_scope))
    smd, gmd = synthesis_metric_tables(params, stage)
    keep = getattr(params, "_w_keep", {})
    keep[stage] = (smd, gmd)
    params._w_keep = keep
    setattr(params, pre + "smd", smd.ctypes.data)
    setattr(params, pre + "gmd", gmd.ctypes.data)


def mass_correct(coef, source, functional, inverse_mats, directions, eps=1e-10) -> np.ndarray:
    """One group's mass correction (poisson.py:166-232): the coefficients
    corrected along the constant-synthesis direction (the outer product of
    the per-axis `directions`) so that the synthesis by `inverse_mats` (one
    per axis, None for identity) carries sum(functional * source)."""
    c = np.ascontiguousarray(coef, dtype=np.float32).copy()
    src = np.ascontiguousarray(source, dtype=np.float32)
    O = np.ascontiguousarray(functional, dtype=np.float64)
    dims = np.asarray(c.shape, dtype=np.int32)
    mats = [None if m is None else np.ascontiguousarray(m, dtype=np.float32) for m in inverse_mats]
    dirs = [np.ascontiguousarray(d, dtype=np.float64) for d in directions]
    mp = (ctypes.c_void_p * len(mats))(*[None if m is None else m.ctypes.data for m in mats])
    dp = (ctypes.c_void_p * len(dirs))(*[d.ctypes.data for d in dirs])
    lib().oracle_mass_correct(c.ndim, _p(dims), mp, dp, _p(O), _p(src), _p(c), eps)
    return c


MASS_MODES = {"none": 0, "ht": 1, "wiener": 2, "both": 3}


def set_mass_conservation(params, mode: str) -> None:
    """poisson_group_mass_conservation with each stage's constant-synthesis
    directions: per axis the solution of M d = 1 for its synthesis matrix M
    in float64 (poisson.py:120-160), the member axis for every n."""
    params.mass_mode = MASS_MODES[mode]
    keep = {}
    for stage in ("ht", "wiener"):
        _, inv, _ = _stage_synthesis(params, stage)
        max_m = params.ht_max if stage == "ht" else params.wi_max
        mdir = np.concatenate([np.linalg.solve(group_inverse(n).astype(np.float64), np.ones(n))
                               for n in range(1, max_m + 1)])
        sdir = np.concatenate([np.linalg.solve(m.astype(np.float64), np.ones(m.shape[0]))
                               for m in inv])
        keep[stage] = (mdir, sdir)
        pre = "ht_" if stage == "ht" else "wi_"
        setattr(params, pre + "mdir", mdir.ctypes.data)
        setattr(params, pre + "sdir", sdir.ctypes.data)
    params._mass_keep = keep


def set_ht_thresholds(params, thresholds: np.ndarray) -> None:
    """Point the oracle's stage-1 matching at per-reference thresholds
    (kept alive on the params object)."""
    arr = np.ascontiguousarray(thresholds, dtype=np.float32)
    params._thr_keep = arr
    params.ht_thresholds = arr.ctypes.data


def set_ht_standardized(params, max_count: int = 512) -> None:
    """Stage-1 matching with the candidate-standardized form of
    params.distance, its null tables cast to float32 as the reference does
    (blockmatching.py:288-319)."""
    m, v = poisson_null_moments(max_count, params.distance or DIST_DEVIANCE)
    params._null_keep = (np.asarray(m, dtype=np.float32), np.asarray(v, dtype=np.float32))
    params.ht_standardize = 1
    params.null_means = params._null_keep[0].ctypes.data
    params.null_vars = params._null_keep[1].ctypes.data
    params.null_max = int(max_count)


# ---- the stationary Gaussian noise model ------------------------------------

def blockmatch_ssd(vol, block, radius, refs, threshold, max_matches, min_matches, *,
                   noise_ssd=None, gamma=0.0, pow2=True, nthreads=0):
    """SSD block matching with the optional Gaussian noise-bias correction
    (blockmatching.py:1046-1205): (sel (nrefs, max) padded with -1, counts,
    accepted) for the patch-grid references refs (nrefs, ndim)."""
    v = np.ascontiguousarray(vol, dtype=np.float32)
    nd = v.ndim
    shp = np.asarray(v.shape, dtype=np.int64)
    blk = np.asarray(block, dtype=np.int32)
    cts = np.asarray([v.shape[d] - block[d] + 1 for d in range(nd)], dtype=np.int64)
    rad = np.asarray(radius, dtype=np.int32)
    r = np.ascontiguousarray(refs, dtype=np.int64).reshape(-1, nd)
    thr = np.asarray(threshold, dtype=np.float32)
    thr = np.ascontiguousarray(np.broadcast_to(thr, (len(r),)), dtype=np.float32)
    ns = None if noise_ssd is None else np.ascontiguousarray(noise_ssd, dtype=np.float32)
    sel = np.empty((len(r), max_matches), np.int64)
    cnt = np.empty(len(r), np.int32)
    acc = np.empty(len(r), np.int64)
    lib().oracle_blockmatch_ssd(_p(v), nd, _p(shp), _p(cts), _p(blk), _p(r), len(r), _p(rad),
                                _p(thr), max_matches, min_matches, int(pow2), _p(ns),
                                float(gamma), nthreads, _p(sel), _p(cnt), _p(acc))
    return sel, cnt, acc


def blockmatch_ssd_poisson(vol, block, radius, refs, threshold, max_matches, min_matches,
                           gamma, count_scale=1.0, pow2=True, nthreads=0):
    """Stage-1 SSD block matching of the Poisson model with the Poisson gamma
    correction (blockmatching.py:432-443, 1046-1205): (sel (nrefs, max) padded
    with -1, counts, accepted) for the patch-grid references refs (nrefs, ndim)."""
    v = np.ascontiguousarray(vol, dtype=np.float32)
    nd = v.ndim
    shp = np.asarray(v.shape, dtype=np.int64)
    blk = np.asarray(block, dtype=np.int32)
    cts = np.asarray([v.shape[d] - block[d] + 1 for d in range(nd)], dtype=np.int64)
    rad = np.asarray(radius, dtype=np.int32)
    r = np.ascontiguousarray(refs, dtype=np.int64).reshape(-1, nd)
    thr = np.asarray(threshold, dtype=np.float32)
    thr = np.ascontiguousarray(np.broadcast_to(thr, (len(r),)), dtype=np.float32)
    sel = np.empty((len(r), max_matches), np.int64)
    cnt = np.empty(len(r), np.int32)
    acc = np.empty(len(r), np.int64)
    lib().oracle_blockmatch_ssd_poisson(_p(v), nd, _p(shp), _p(cts), _p(blk), _p(r), len(r),
                                        _p(rad), _p(thr), max_matches, min_matches, int(pow2),
                                        float(np.float32(gamma)), float(count_scale), nthreads,
                                        _p(sel), _p(cnt), _p(acc))
    return sel, cnt, acc


def set_poisson_ssd(params, *, ht_match, wi_match, gamma) -> None:
    """Switch the Poisson params to the SSD distance: the profile's match
    thresholds (they scale with the matched volume's squared range) and the
    stage-1 Poisson gamma (0: plain SSD)."""
    params.poisson_ssd = 1
    params.ht_gamma = float(np.float32(gamma))
    params.ht_match_d, params.wi_match_d = float(ht_match), float(wi_match)


def gaussian_group_variance(maps, positions, group_matrix, k: int) -> np.ndarray:
    """The (n, P) noise variance of one group under the stationary model
    (variance.py:1889-2106): maps (P,) + domain float32, positions (n, ndim)
    patch origins, the n x n member analysis matrix."""
    m = np.ascontiguousarray(maps, dtype=np.float32)
    P = m.shape[0]
    domain = m.shape[1:]
    pos = np.asarray(positions, dtype=np.int64)
    n = pos.shape[0]
    lag = np.zeros((n, n), dtype=np.int64)
    for a, d in enumerate(domain):
        lag = lag * d + (pos[:, None, a] - pos[None, :, a]) % d
    g = np.ascontiguousarray(group_matrix, dtype=np.float32)
    out = np.empty((n, P), np.float32)
    lib().oracle_gaussian_group_variance(_p(m), P, int(np.prod(domain)), _p(lag), _p(g), n,
                                         int(k), _p(out))
    return out


def ht_gaussian(coef, sigma, lam=3.0):
    c = np.ascontiguousarray(coef, dtype=np.float32).copy()
    s = np.ascontiguousarray(sigma, dtype=np.float32)
    att = np.empty_like(c)
    lib().oracle_ht_gaussian(_p(c), _p(s), c.size, lam, _p(att))
    return c, att


def wiener_gaussian(coef, ref_coef, sigma, mu=0.4):
    c = np.ascontiguousarray(coef, dtype=np.float32).copy()
    r = np.ascontiguousarray(ref_coef, dtype=np.float32)
    s = np.ascontiguousarray(sigma, dtype=np.float32)
    att = np.empty_like(c)
    sp = np.empty_like(c)
    lib().oracle_wiener_gaussian(_p(c), _p(r), _p(s), c.size, mu, _p(att), _p(sp))
    return c, att, sp


def set_gaussian(params, *, global_sigma, domain, ht_maps, wi_maps, ht_match, wi_match,
                 ht_noise_ssd=None, gamma=0.0, k=None) -> None:
    """Switch the params to the stationary Gaussian model: the tables of
    each stage (float32 cross-covariance maps (P,) + domain, the HT search-box
    noise SSD or None), the match thresholds as given (they scale with the
    matched volume's squared range) and k (None keeps the params' k)."""
    params.gaussian = 1
    params.global_sigma = float(global_sigma)
    for d, v in enumerate(domain):
        params.cov_domain[d] = int(v)
    keep = [np.ascontiguousarray(ht_maps, np.float32), np.ascontiguousarray(wi_maps, np.float32)]
    params.ht_cov, params.wi_cov = keep[0].ctypes.data, keep[1].ctypes.data
    if ht_noise_ssd is not None:
        keep.append(np.ascontiguousarray(ht_noise_ssd, np.float32))
        params.ht_noise_ssd = keep[-1].ctypes.data
    else:
        params.ht_noise_ssd = None
    params.ht_gamma = float(gamma)
    params.ht_match_d, params.wi_match_d = float(ht_match), float(wi_match)
    if k is not None:
        params.k = int(k)
    params._gauss_keep = keep


def bmnd_gaussian_sigma(vol, params, *, capture_matches: bool = False):
    """The oracle's two-stage Gaussian pipeline with the sigma maps:
    (out, pilot, ht_map, wi_map[, (ht_sel, ht_cnt), (wi_sel, wi_cnt)])."""
    v = np.ascontiguousarray(vol, dtype=np.float32)
    if not params.gaussian:
        raise ValueError("set_gaussian() first")
    out, pilot = np.empty_like(v), np.empty_like(v)
    hm, wm = np.empty_like(v), np.empty_like(v)
    nd = v.ndim
    cnt_ht = tuple(int(v.shape[d]) - params.ht_block[d] + 1 for d in range(nd))
    cnt_wi = tuple(int(v.shape[d]) - params.wi_block[d] + 1 for d in range(nd))
    n_ht = len(schedule(cnt_ht, tuple(params.ht_block[:nd]), tuple(params.ht_step[:nd]),
                        params.shift_density, params.schedule_density))
    n_wi = len(schedule(cnt_wi, tuple(params.wi_block[:nd]), tuple(params.wi_step[:nd]),
                        params.shift_density, params.schedule_density))
    ht = (np.empty((n_ht, params.ht_max), np.int64), np.empty(n_ht, np.int32))
    wi = (np.empty((n_wi, params.wi_max), np.int64), np.empty(n_wi, np.int32))
    g = ctypes.c_double(0.0)
    rc = lib().oracle_bmnd_poisson_sigma(_p(v), ctypes.byref(params), _p(out), _p(pilot),
                                         _p(ht[0]), _p(ht[1]), n_ht, _p(wi[0]), _p(wi[1]), n_wi,
                                         _p(hm), _p(wm), ctypes.addressof(g))
    if rc != 0:
        raise RuntimeError(f"oracle Gaussian pipeline failed with code {rc}")
    if capture_matches:
        return out, pilot, hm, wm, ht, wi
    return out, pilot, hm, wm


def bmnd_gaussian(vol, params, *, capture_matches: bool = False):
    """(out, pilot[, (ht_sel, ht_cnt), (wi_sel, wi_cnt)]) of the Gaussian
    pipeline."""
    v = np.ascontiguousarray(vol, dtype=np.float32)
    if not params.gaussian:
        raise ValueError("set_gaussian() first")
    out, pilot = np.empty_like(v), np.empty_like(v)
    args = [None, None, 0, None, None, 0]
    ht = wi = None
    if capture_matches:
        nd = v.ndim
        cnt_ht = tuple(int(v.shape[d]) - params.ht_block[d] + 1 for d in range(nd))
        cnt_wi = tuple(int(v.shape[d]) - params.wi_block[d] + 1 for d in range(nd))
        n_ht = len(schedule(cnt_ht, tuple(params.ht_block[:nd]), tuple(params.ht_step[:nd]),
                            params.shift_density, params.schedule_density))
        n_wi = len(schedule(cnt_wi, tuple(params.wi_block[:nd]), tuple(params.wi_step[:nd]),
                            params.shift_density, params.schedule_density))
        ht = (np.empty((n_ht, params.ht_max), np.int64), np.empty(n_ht, np.int32))
        wi = (np.empty((n_wi, params.wi_max), np.int64), np.empty(n_wi, np.int32))
        args = [ht[0], ht[1], n_ht, wi[0], wi[1], n_wi]
    rc = lib().oracle_bmnd_poisson(_p(v), ctypes.byref(params), _p(out), _p(pilot),
                                   _p(args[0]), _p(args[1]), args[2], _p(args[3]),
                                   _p(args[4]), args[5])
    if rc != 0:
        raise RuntimeError(f"oracle Gaussian pipeline failed with code {rc}")
    if capture_matches:
        return out, pilot, ht, wi
    return out, pilot
