"""The exact planes by member pairs (bmnd_hip_exact_std_pairs, the Wiener
stage's default order) against the reference's serial order.

The pair form sums the same quantity in another order, so the bar is float32
rounding, not bit equality: per group and plane, every coefficient std
within 1e-5 of that plane's largest std from the serial kernel (which equals
the oracle bit for bit), and both within 1e-5 of a float64 evaluation.  End
to end (default order) the match lists of both stages and the pilot stay
bit-identical to the oracle, and the output is held to the 1e-4 max relative
error and 0.01 dB PSNR bar of the float32 path."""

import ctypes

import numpy as np
import pytest
import torch

import oracle
from bmnd_amd import BM3DProfile, BM4DProfile, NoiseModel, _lib, bmnd, engine, phantoms
from bmnd_amd.engine import PoissonPipeline, _ptr, make_plan
from bmnd_amd.transforms import group_matrices, spatial_matrices

pytestmark = [pytest.mark.gpu, pytest.mark.default_exact_order]
DEV = torch.device("cuda", 0)


def _groups(shape, block, max_m, sizes, spread, seed):
    """Random groups: for each size n, members at distinct patch origins
    within +-spread of a random centre (heavily overlapping for small
    spread); (sel, counts) as the kernels take them."""
    rng = np.random.default_rng(seed)
    grid = [shape[d] - block[d] + 1 for d in range(3)]
    sel = np.full((len(sizes), max_m), -1, np.int32)
    for r, n in enumerate(sizes):
        c = [int(rng.integers(0, grid[d])) for d in range(3)]
        seen = set()
        while len(seen) < n:
            o = tuple(int(np.clip(c[d] + rng.integers(-spread[d], spread[d] + 1), 0, grid[d] - 1))
                      for d in range(3))
            seen.add(o)
        lins = [(o[0] * grid[1] + o[1]) * grid[2] + o[2] for o in sorted(seen)]
        rng.shuffle(lins)
        sel[r, :n] = lins
    return sel, np.asarray(sizes, np.int32), grid


def _run(entry, var, shape, plan, sel, cnt, k):
    L = _lib.lib()
    R = len(cnt)
    kcap = min(k, plan.max_m)
    out = torch.full((R * kcap * plan.P,), float("nan"), dtype=torch.float32, device=DEV)
    mats, _ = plan.device_constants(DEV)
    s = ctypes.c_void_p(torch.cuda.current_stream(DEV).cuda_stream)
    sel_d = torch.from_numpy(sel.ravel()).to(DEV)
    cnt_d = torch.from_numpy(cnt).to(DEV)
    _lib.check(getattr(L, entry)(_ptr(var), (ctypes.c_int64 * 3)(*shape),
                                 (ctypes.c_int32 * 3)(*plan.block3),
                                 (ctypes.c_int32 * 3)(*plan.radius3), _ptr(sel_d), _ptr(cnt_d),
                                 R, plan.max_m, _ptr(mats), k, _ptr(out), s))
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(R, kcap, plan.P)


def _f64_planes(vv, pos, block, g, spatial, planes):
    """sum_v var_v (sum_(i,p) G[q,i] S[s,p])^2 in float64 over union voxels."""
    S = np.array([[1.0]])
    for m in spatial:
        S = np.kron(S, np.asarray(m, np.float64))
    P = S.shape[0]
    local = np.stack(np.meshgrid(*[np.arange(b) for b in block], indexing="ij"), -1).reshape(P, 3)
    acc = {}
    for i in range(pos.shape[0]):
        for p in range(P):
            acc.setdefault(tuple(pos[i] + local[p]), []).append((i, p))
    out = np.zeros((planes, P))
    g64 = np.asarray(g, np.float64)[:planes]
    for x, ents in acc.items():
        a = np.zeros((planes, P))
        for i, p in ents:
            a += np.outer(g64[:, i], S[:, p])
        out += float(vv[x]) * a * a
    return np.sqrt(out)


@pytest.mark.parametrize("geom", ["wiener-5", "ht-4", "image-8"])
@pytest.mark.parametrize("k", [4, 8])
def test_pair_planes_match_serial_planes(geom, k):
    """Every group size up to the stack, overlapping and scattered members,
    k = 8 exercising a second round of four planes."""
    print(f"exact pairs: {geom}, k={k}", flush=True)
    if geom == "image-8":
        prof, stage, shape, spread = BM3DProfile(noise_model=NoiseModel.POISSON), "wiener", \
            (1, 70, 66), (0, 6, 6)
    else:
        prof = BM4DProfile(noise_model=NoiseModel.POISSON)
        stage = "wiener" if geom == "wiener-5" else "ht"
        shape, spread = (26, 30, 28), (3, 3, 3)
    plan = make_plan(prof, stage, 2 if geom == "image-8" else 3)
    rng = np.random.default_rng(11)
    vol = rng.uniform(0.05, 20.0, size=shape).astype(np.float32)
    var = torch.from_numpy(vol).to(DEV)
    sizes = [1, 2, 3, 4, 5, 7, 8, 11, 16] + ([31, 32] if plan.max_m >= 32 else []) * 3
    sizes = [n for n in sizes if n <= plan.max_m] * 4
    sel, cnt, grid = _groups(shape, plan.block3, plan.max_m, sizes, spread, seed=5)
    # scattered members too (few or no overlaps), still inside one search box
    # of the centre, as every group of a stage is
    sel2, cnt2, _ = _groups(shape, plan.block3, plan.max_m, sizes[: len(sizes) // 2],
                            plan.radius3, seed=6)
    sel, cnt = np.concatenate([sel, sel2]), np.concatenate([cnt, cnt2])
    ser = _run("bmnd_hip_exact_std", var, shape, plan, sel, cnt, k)
    par = _run("bmnd_hip_exact_std_pairs", var, shape, plan, sel, cnt, k)
    worst = 0.0
    for r in range(len(cnt)):
        q = min(k, int(cnt[r]))
        a, b = par[r, :q], ser[r, :q]
        assert np.all(np.isfinite(a)), (r, cnt[r])
        scale = np.maximum(b.max(axis=1, keepdims=True), 1e-30)
        err = float(np.max(np.abs(a - b) / scale))
        worst = max(worst, err)
    print(f"  {len(cnt)} groups, worst |pairs - serial| / plane max = {worst:.3g}", flush=True)
    assert worst <= 1e-5
    # both orders against float64 on a few groups
    g_mats = {}
    spatial = [np.ones((1, 1)) if b == 1 else spatial_matrices(plan.spatial, b)[0]
               for b in plan.block3]
    for r in range(0, len(cnt), 7):
        n = int(cnt[r])
        q = min(k, n)
        pos = np.stack(np.unravel_index(sel[r, :n].astype(np.int64), grid), axis=1)
        if n not in g_mats:
            g_mats[n] = group_matrices(n)[0]
        want = _f64_planes(vol, pos, plan.block3, g_mats[n], spatial, q)
        scale = np.maximum(want.max(axis=1, keepdims=True), 1e-30)
        assert np.max(np.abs(par[r, :q] - want) / scale) <= 1e-5
        assert np.max(np.abs(ser[r, :q] - want) / scale) <= 1e-5


def _oracle_compare(noisy, prof, params):
    pipe = PoissonPipeline(prof, noisy.shape, DEV)
    ht, wi = pipe.run(torch.from_numpy(noisy).to(DEV))
    torch.cuda.synchronize()
    o_out, o_pilot, (o_hsel, o_hcnt), (o_wsel, o_wcnt) = oracle.bmnd_poisson(
        noisy, params, capture_matches=True)
    for g, os_, oc in ((ht, o_hsel, o_hcnt), (wi, o_wsel, o_wcnt)):
        gs, gc = g.sel.cpu().numpy(), g.counts.cpu().numpy()
        assert np.array_equal(gc, oc)
        for r in range(gs.shape[0]):
            assert np.array_equal(gs[r, : gc[r]], os_[r, : oc[r]])
    pilot, out = ht.est.cpu().numpy(), wi.est.cpu().numpy()
    assert pilot.tobytes() == o_pilot.tobytes()
    rel = float(np.max(np.abs(out - o_out)) / np.max(np.abs(o_out)))
    return out, o_out, rel


@pytest.mark.parametrize("case", ["readme-64", "vol-28x32x24", "vol-40", "blob-144"])
def test_default_order_pipeline_within_tolerance(case):
    assert engine.EXACT_ORDER == {"ht": "serial", "wiener": "pairs"}
    assert engine.WIENER_TRANSFORMS == ["mfma"]
    print(f"default exact order end to end: {case}", flush=True)
    if case == "readme-64":
        truth, noisy = phantoms.square_image(0)
        prof, nd = BM3DProfile(noise_model=NoiseModel.POISSON), 2
    elif case == "blob-144":
        truth, noisy = phantoms.blob_image(144, seed=7, peak=20.0)
        prof, nd = BM3DProfile(noise_model=NoiseModel.POISSON), 2
    else:
        shape = (28, 32, 24) if case == "vol-28x32x24" else (40, 40, 40)
        truth, noisy = phantoms.ellipsoid_volume(shape, seed=3, peak=10.0, n_ellipsoids=8,
                                                 axes_range=(3.0, 14.0))
        prof, nd = BM4DProfile(noise_model=NoiseModel.POISSON), 3
    out, o_out, rel = _oracle_compare(noisy, prof, oracle.default_params(nd, noisy.shape))
    dpsnr = abs(phantoms.psnr(out, truth) - phantoms.psnr(o_out, truth))
    bits = int(np.sum(out.view(np.int32) != o_out.view(np.int32)))
    print(f"  output max rel {rel:.3g}, |dPSNR| {dpsnr:.2g} dB, {bits}/{out.size} values differ",
          flush=True)
    assert rel <= 1e-4
    assert dpsnr <= 0.01


def test_default_order_deterministic_graph_and_batch():
    """The pair order is deterministic: two calls, the captured graph and a
    batch give the same bits."""
    print("default exact order: determinism, graph, batch", flush=True)
    _, a = phantoms.ellipsoid_volume((30, 34, 28), seed=21, peak=8.0, n_ellipsoids=8,
                                     axes_range=(3.0, 12.0))
    _, b = phantoms.ellipsoid_volume((24, 26, 30), seed=22, peak=20.0, n_ellipsoids=6,
                                     axes_range=(3.0, 10.0))
    prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    pipe = PoissonPipeline(prof, a.shape, DEV)
    va = torch.from_numpy(a).to(DEV)
    _, w1 = pipe.run(va)
    _, w2 = pipe.run(va)
    assert torch.equal(w1.est.view(torch.int32), w2.est.view(torch.int32))
    g = pipe.run_graph(va).clone()
    assert torch.equal(g.view(torch.int32), w1.est.view(torch.int32))
    from bmnd_amd import bmnd_batch
    outs = bmnd_batch([a, b], prof)
    assert outs[0].tobytes() == w1.est.cpu().numpy().tobytes()
    pb = PoissonPipeline(prof, b.shape, DEV)
    _, wb = pb.run(torch.from_numpy(b).to(DEV))
    assert outs[1].tobytes() == wb.est.cpu().numpy().tobytes()
    # switching the order re-captures the graph
    prev = engine.set_exact_order(wiener="serial")
    try:
        gs = pipe.run_graph(va).clone()
        _, ws = pipe.run(va)
        assert torch.equal(gs.view(torch.int32), ws.est.view(torch.int32))
    finally:
        engine.set_exact_order(**prev)


def test_stage_driver_pair_flag():
    """bmnd_hip_stage with BMND_HIP_EXACT_PAIRS_WIENER in k equals the
    call-by-call Wiener stage in the pair order."""
    print("stage driver: pair flag", flush=True)
    _, noisy = phantoms.ellipsoid_volume((32, 30, 34), seed=31, peak=10.0, n_ellipsoids=6,
                                         axes_range=(3.0, 10.0))
    prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    pipe = PoissonPipeline(prof, noisy.shape, DEV)
    v = torch.from_numpy(noisy).to(DEV)
    ht, wi = pipe.run(v)
    L = _lib.lib()
    plan = pipe.plans["wiener"]
    shape = (ctypes.c_int64 * 3)(*pipe.shape3)
    block = (ctypes.c_int32 * 3)(*plan.block3)
    radius = (ctypes.c_int32 * 3)(*plan.radius3)
    refs = pipe._ref_lin(plan)
    R = refs.numel()
    k = engine.k_with_order(pipe.k)
    assert k == pipe.k | (1 << 17) | (1 << 18)
    need = L.bmnd_hip_stage_workspace(shape, block, R, plan.max_m, k)
    assert need == L.bmnd_hip_stage_workspace(shape, block, R, plan.max_m, pipe.k)
    ws = torch.empty(need, dtype=torch.uint8, device=DEV)
    est = torch.empty_like(v)
    sel = torch.empty(R * plan.max_m, dtype=torch.int32, device=DEV)
    cnt = torch.empty(R, dtype=torch.int32, device=DEV)
    acc = torch.empty(R, dtype=torch.int32, device=DEV)
    mats, kai = plan.device_constants(DEV)
    var2 = pipe.variance(ht.est)
    stream = ctypes.c_void_p(torch.cuda.current_stream(DEV).cuda_stream)
    _lib.check(L.bmnd_hip_stage(1, _ptr(ht.est), _ptr(v), _ptr(var2), _ptr(v), shape, block,
                                radius, _ptr(refs), R, plan.threshold, plan.max_m, plan.min_m,
                                pipe.count_scale, -1, _ptr(mats), _ptr(kai), plan.shrink, k,
                                _ptr(ws), need, _ptr(est), _ptr(sel), _ptr(cnt), _ptr(acc),
                                stream))
    torch.cuda.synchronize()
    assert torch.equal(sel.view(R, plan.max_m), wi.sel) and torch.equal(cnt, wi.counts)
    assert torch.equal(est.view(torch.int32), wi.est.view(torch.int32))


def test_bmnd_default_order_against_serial():
    """bmnd() in the default order against bmnd() in the serial order on a
    larger volume: same lists (through the same pilot), output within the
    float32 bar."""
    print("bmnd default vs serial order, 64^3", flush=True)
    truth, noisy = phantoms.ellipsoid_volume((64, 64, 64), seed=9, peak=10.0, n_ellipsoids=12,
                                             axes_range=(4.0, 20.0))
    prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    out_p = bmnd(noisy, prof)
    prev = engine.set_exact_order(wiener="serial", transforms="serial")
    try:
        out_s = bmnd(noisy, prof)
    finally:
        engine.set_exact_order(**prev)
    rel = float(np.max(np.abs(out_p - out_s)) / np.max(np.abs(out_s)))
    dpsnr = abs(phantoms.psnr(out_p, truth) - phantoms.psnr(out_s, truth))
    print(f"  max rel {rel:.3g}, |dPSNR| {dpsnr:.2g} dB", flush=True)
    assert rel <= 1e-4 and dpsnr <= 0.01
