#!/usr/bin/env python3
"""Throughput of the Poisson bmnd() pipeline on one MI355X.

One step = the full two-stage bmnd() (HT pilot + Wiener) over one synthetic
input that is already resident in HBM.  Prints ONE JSON line:
  value        input voxels per second over the timed steps
  roofline     the dominant kernel against its MI355X roof (HIP events)
  cpu_baseline the CPU oracle timed on a bounded sample of the same workload

Workloads (--config):
  bm4d-256     256^3 volume of 40 random ellipsoids + 0.05 floor, peak 10
               counts, BM4DProfile(noise_model=POISSON) defaults   [default]
  bm3d-512x16  16 images of 512^2 (Gaussian blobs + rectangles, peaks in
               {1,5,20,100}), BM3DProfile(noise_model=POISSON) defaults
  bm3d-64      64^2 README square image, BM3DProfile defaults
  batch-uneven 32 images with sides uniform in [96, 1024] and 16 volumes
               with sides uniform in [48, 128], peaks uniform in [1, 50]:
               one batch per dimensionality (BM3D / BM4D defaults), every
               kernel launched once per batch over all its inputs
  gauss-bm4d-256     the stationary Gaussian model: the 256^3 ellipsoid
               phantom scaled to [0, 1] plus white noise of std 0.1,
               BM4DProfile() defaults (nf 16^3, k 4, gamma 3), sigma given
  gauss-bm3d-512x16  16 blob images of 512^2 scaled to [0, 1] plus white
               noise of std 0.1, BM3DProfile() defaults (nf 32^2), sigma given
With --batch, the inputs of a multi-input workload (bm3d-512x16) run as one
batch instead of one bmnd() call after another.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "voxels/s of bmnd() on 1 MI355X, plain vs surrogate pilot, at fixed PSNR gap"
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# block matching's roof: the reference's 2 logarithms + 8 arithmetic
# operations per deviance term, at the transcendental rate (quarter rate of
# the 64-lane issue, 256 CUs x 16 x 2.4 GHz) and the non-packed FP32 VALU rate
R_TRANS = 9.8e12
R_VALU = 3.9e13
TERM_RATE = 1.0 / max(2.0 / R_TRANS, 8.0 / R_VALU)  # deviance terms per second
# workgroup size of the block-matching kernel per deviance mode (its last
# template argument, part of the kernel name rocprofv3 reports)
BM_THREADS = {0: 512, 1: 1024, 2: 512, 3: 512, 4: 512}
# stack capacity of the exact-variance kernel per block
EXACT_CONFIG = {(1, 8, 8): 32, (4, 4, 4): 16, (5, 5, 5): 32}


def parse() -> argparse.Namespace:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="bm4d-256",
                    choices=["bm4d-256", "bm3d-512x16", "bm3d-64", "batch-uneven",
                             "gauss-bm4d-256", "gauss-bm3d-512x16"])
    ap.add_argument("--batch", action="store_true",
                    help="run the inputs of the workload as one batch (bmnd_batch)")
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="side of the cube (3-D) or square (2-D) the CPU oracle is timed on "
                         "(default 64 for volumes, 256 for images)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the learned-pilot leg (3-D workloads)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB mebibytes (for scaling "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE readings) and exit")
    ap.add_argument("--deviance", default="auto-fast",
                    choices=["auto", "exact", "one-log", "auto-one-log", "certified",
                             "auto-certified", "fast", "auto-fast"],
                    help="block-matching deviance evaluation (integer counts use the "
                         "term table; other inputs the certified one-log form, whose keys "
                         "equal the reference expression's, or with 'auto' the reference's "
                         "two-log form)")
    return ap.parse_args()


def make_inputs(config: str):
    from bmnd_amd import BM3DProfile, BM4DProfile, NoiseModel, phantoms
    if config == "bm4d-256":
        truth, noisy = phantoms.ellipsoid_volume((256, 256, 256), seed=0, peak=10.0)
        return BM4DProfile(noise_model=NoiseModel.POISSON), [noisy], [truth], {
            "workload": "256^3 BM4D Poisson, 40 ellipsoids + 0.05 floor, peak 10 counts",
            "shape": [256, 256, 256], "profile": "BM4DProfile(noise_model=POISSON)", "seed": 0}
    if config == "bm3d-512x16":
        pairs = [phantoms.blob_image(512, seed=s) for s in range(16)]
        return BM3DProfile(noise_model=NoiseModel.POISSON), [p[1] for p in pairs], \
            [p[0] for p in pairs], {
                "workload": "16 x 512^2 BM3D Poisson images, Gaussian blobs + rectangles, "
                            "peaks {1,5,20,100}",
                "shape": [16, 512, 512], "profile": "BM3DProfile(noise_model=POISSON)",
                "seeds": list(range(16))}
    truth, noisy = phantoms.square_image(0)
    return BM3DProfile(noise_model=NoiseModel.POISSON), [noisy], [truth], {
        "workload": "64^2 README square image, BM3D Poisson", "shape": [64, 64],
        "profile": "BM3DProfile(noise_model=POISSON)", "seed": 0}


def measured_traffic(config: str, deviance: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_traffic.json")
    try:
        data = json.load(open(path))
        k = data[f"{config}/{deviance}"]["kernels"][kernel]
        return int(k["fetch_bytes"] + k["write_bytes"])
    except (OSError, KeyError, ValueError):
        return None


def crop(a: np.ndarray, side: int) -> np.ndarray:
    c = [(s - side) // 2 for s in a.shape]
    return np.ascontiguousarray(a[tuple(slice(c[d], c[d] + side) for d in range(a.ndim))])


def fp64_check(profile, noisy: np.ndarray, truth: np.ndarray, side: int) -> dict:
    """On the CPU-baseline sample: the GPU output against the float32 oracle
    (max relative error and bits; and bits with both stages' exact planes in
    the reference's serial order, which the oracle follows) and against the
    oracle's float64 run (the error of the float32 path), PSNR against the
    noiseless phantom."""
    import torch

    import oracle
    from bmnd_amd import engine
    from bmnd_amd.engine import PoissonPipeline
    from bmnd_amd.phantoms import psnr
    side = min(side or (64 if noisy.ndim == 3 else 256), *noisy.shape)
    sample, tsample = crop(noisy, side), crop(truth, side)
    dev = torch.device("cuda", 0)
    pipe = PoissonPipeline(profile, sample.shape, dev)
    v = torch.from_numpy(sample).to(dev)
    gpu = pipe.run(v)[1].est.cpu().numpy()
    prev = engine.set_exact_order(ht="serial", wiener="serial", transforms="serial")
    try:
        gpu_serial = pipe.run(v)[1].est.cpu().numpy()
    finally:
        engine.set_exact_order(**prev)
    threads = host_threads()
    o32, _ = oracle.bmnd_poisson(sample, nthreads=threads)
    o64, _ = oracle.bmnd_poisson_f64(sample, nthreads=threads)
    scale = float(np.max(np.abs(o32)))
    return {"sample": f"{side}^{sample.ndim} centre crop",
            "exact_order": dict(engine.EXACT_ORDER, wiener_transforms=engine.WIENER_TRANSFORMS[0]),
            "gpu_vs_oracle_max_rel": float(np.max(np.abs(gpu - o32)) / scale),
            "gpu_vs_oracle_bits_differing": int(np.sum(gpu.view(np.int32) != o32.view(np.int32))),
            "gpu_serial_order_vs_oracle_bits_differing":
                int(np.sum(gpu_serial.view(np.int32) != o32.view(np.int32))),
            "gpu_vs_fp64_max_rel": float(np.max(np.abs(gpu - o64)) / np.max(np.abs(o64))),
            "oracle_vs_fp64_max_rel": float(np.max(np.abs(o32 - o64)) / np.max(np.abs(o64))),
            "psnr_gpu_db": round(psnr(gpu, tsample), 4),
            "psnr_oracle_db": round(psnr(o32, tsample), 4),
            "psnr_fp64_db": round(psnr(o64, tsample), 4)}


def host_threads() -> int:
    """The host cores this process may use, at most the 16 a one-GPU share
    of the machine provides."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def gpu_on_sample(profile, sample: np.ndarray, reps: int = 3) -> dict:
    """The GPU timed on the CPU leg's own sample (warm, input in HBM)."""
    import torch

    from bmnd_amd.engine import PoissonPipeline
    dev = torch.device("cuda", 0)
    pipe = PoissonPipeline(profile, sample.shape, dev)
    v = torch.from_numpy(np.ascontiguousarray(sample)).to(dev)
    pipe.run(v)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        pipe.run(v)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    return {"value": float(sample.size / dt), "ms": round(1e3 * dt, 3)}


def cpu_baseline(config: str, noisy: np.ndarray, side: int) -> dict:
    import oracle
    threads = host_threads()
    side = min(side or (64 if noisy.ndim == 3 else 256), *noisy.shape)
    if noisy.ndim == 3:
        c = [(s - side) // 2 for s in noisy.shape]
        sample = np.ascontiguousarray(noisy[c[0]:c[0] + side, c[1]:c[1] + side, c[2]:c[2] + side])
        desc = f"{side}^3 centre crop of the benchmark volume, oracle bmnd_poisson (both stages)"
    else:
        c = [(s - side) // 2 for s in noisy.shape]
        sample = np.ascontiguousarray(noisy[c[0]:c[0] + side, c[1]:c[1] + side])
        desc = f"{side}^2 centre crop of the first image, oracle bmnd_poisson (both stages)"
    oracle.build()
    t0 = time.perf_counter()
    oracle.bmnd_poisson(sample, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": float(sample.size / dt), "unit": "voxels/s", "cores": threads,
            "kind": "port", "sample": desc, "seconds": round(dt, 3), "_sample": sample}


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(every run of this build uses a single MI355X).", file=sys.stderr)
        return 2
    import torch

    if args.calibrate_copy:
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4}))
        return 0

    from bmnd_amd.engine import (EXACT_ORDER, KernelTimer, PoissonPipeline, blockmatch_work,
                                 exact_std_work)

    if args.batch or args.config == "batch-uneven":
        return run_batch(args)
    if args.config.startswith("gauss-"):
        return run_gaussian(args)
    profile, noisy_list, truth_list, cfg = make_inputs(args.config)
    dev = torch.device("cuda", 0)
    vols = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in noisy_list]
    voxels = sum(int(v.numel()) for v in vols)
    # cold call: a pipeline for a shape not seen before (schedules built on
    # the device, tables and plans uploaded) and its first two-stage call
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pipe = PoissonPipeline(profile, noisy_list[0].shape, dev, deviance=args.deviance)
    pipe.run(vols[0])
    torch.cuda.synchronize()
    cold_ms = 1e3 * (time.perf_counter() - t0)

    last_outs = {}

    def step(timer=None):
        outs = []
        for v in vols:
            ht, wi = pipe.run(v, timer)
            last_outs["ht"], last_outs["wiener"] = ht, wi
            outs.append(wi.est)
        return outs

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    timer = KernelTimer()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        outs = step(timer)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps

    # host-to-host: one further call from host arrays, upload and download included
    t0 = time.perf_counter()
    for v_np in noisy_list:
        v_dev = torch.from_numpy(np.ascontiguousarray(v_np)).to(dev)
        _, wi_h = pipe.run(v_dev)
        _ = wi_h.est.cpu().numpy()
    torch.cuda.synchronize()
    h2h_ms = 1e3 * (time.perf_counter() - t0)

    spans = timer.totals_ms()
    per_kernel = {k: float(np.mean(v)) for k, v in spans.items()}
    shape3 = pipe.shape3
    rooflines = {}
    for st in ("ht", "wiener"):
        plan = pipe.plans[st]
        rooflines[f"{st}.blockmatch"] = matcher_roofline(
            plan, pipe.last_modes.get(st, 0), noisy_list[0].shape, per_kernel[f"{st}.blockmatch"],
            measured_traffic(args.config, args.deviance, matcher_kernel(plan, pipe.last_modes.get(st, 0))))
        if f"{st}.exact_std" in per_kernel:
            out = last_outs[st]
            w = exact_std_work(plan, shape3, out.sel, out.counts, pipe.k)
            order = EXACT_ORDER[st]
            rooflines[f"{st}.exact_std"] = exact_roofline(
                plan, w, per_kernel[f"{st}.exact_std"],
                measured_traffic(args.config, args.deviance, exact_kernel(plan, order)), order)
    dominant = max(per_kernel, key=lambda k: per_kernel[k])
    roof = rooflines.get(dominant) or {
        "kernel": dominant, "bound": None, "achieved": None, "peak": None, "unit": None,
        "frac": None, "traffic": None, "avg_ms": round(per_kernel[dominant], 3)}
    work = {s: blockmatch_work(pipe.plans[s], noisy_list[0].shape) for s in ("ht", "wiener")}

    psnr = None
    if truth_list is not None:
        from bmnd_amd.phantoms import psnr as _psnr
        out0 = outs[0].cpu().numpy()
        psnr = {"noisy_db": round(_psnr(noisy_list[0], truth_list[0]), 3),
                "out_db": round(_psnr(out0, truth_list[0]), 3)}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.config, noisy_list[0], args.cpu_sample)
        sample = cpu.pop("_sample")
        # like for like: the GPU on the CPU leg's own sample
        cpu["gpu_on_sample"] = gpu_on_sample(profile, sample)
        cpu["gpu_over_cpu_on_sample"] = round(cpu["gpu_on_sample"]["value"] / cpu["value"], 1)
        # the same leg checks the GPU on the sample against the oracle's runs
        cpu["fp64_check"] = fp64_check(profile, noisy_list[0], truth_list[0], args.cpu_sample)

    value = voxels / (ms_per_step * 1e-3)
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "voxels/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32", "data": "synthetic",
        "config": dict(cfg, pilot="plain", deviance=args.deviance,
                       blockmatch_modes=dict(pipe.last_modes)),
        "roofline": roof, "rooflines": rooflines, "cpu_baseline": cpu,
        "host_to_host": {"ms": round(h2h_ms, 3), "value": round(voxels / (h2h_ms * 1e-3), 1),
                         "unit": "voxels/s"},
        "cold_call_ms": round(cold_ms, 1),
        "kernels_ms": {k: round(v, 3) for k, v in sorted(per_kernel.items())},
        "blockmatch_terms": {s: work[s][1] for s in work},
        "psnr": psnr,
    }
    if len(vols) == 1 and noisy_list[0].size <= (1 << 22):
        line["graph"] = graph_leg(pipe, vols[0], args.steps)
    if noisy_list[0].ndim == 3 and not args.no_surrogate:
        line["surrogate"] = surrogate_leg(profile, noisy_list[0].shape, dev)
    print(json.dumps(line), flush=True)
    return 0


def matcher_kernel(plan, mode) -> str:
    if mode in (3, 4):
        return "k_blockmatch_{}<{}, {}, {}, {}, {}, {}, {}>(BMArgs)".format(
            "cert" if mode == 3 else "fast", *plan.block3, *plan.radius3, BM_THREADS[mode])
    return "k_blockmatch<{}, {}, {}, {}, {}, {}, {}, {}>(BMArgs)".format(
        *plan.block3, *plan.radius3, min(mode, 6) if mode < 5 else (5 if mode < 7 else 6),
        BM_THREADS.get(mode, 512))


def exact_kernel(plan, order: str = "serial") -> str:
    """The name rocprofv3 reports for a stage's exact-plane kernel."""
    nmax = EXACT_CONFIG[tuple(plan.block3)]
    if order == "pairs":
        b0, b1, b2 = plan.block3
        if b0 == b1 == b2:
            return f"k_exact_pairs_rows<{b0}, {nmax}, 4>(ExactArgs)"
        return f"k_exact_pairs<{b0}, {b1}, {b2}, {nmax}, 4>(ExactArgs)"
    return "k_exact_std<{}, {}, {}, {}, {}, {}, {}>(ExactArgs)".format(
        *plan.block3, *plan.radius3, nmax)


def matcher_roofline(plan, mode, shape, ms, traffic) -> dict:
    """Block matching against its transcendental roof: the reference's
    deviance (2 logarithms + 8 FP32-class operations per term) at
    TERM_RATE, and the executed floor of the fast matcher's float pass (one
    hardware log per term) at R_TRANS."""
    from bmnd_amd.engine import blockmatch_work
    _, terms = blockmatch_work(plan, shape)
    ach = terms / (ms * 1e-3) / 1e12
    return {"kernel": matcher_kernel(plan, mode), "bound": "transcendental",
            "achieved": round(ach, 4), "peak": round(TERM_RATE / 1e12, 4),
            "unit": "Tterm/s", "frac": round(ach * 1e12 / TERM_RATE, 4),
            "one_log_floor_frac": round(terms / R_TRANS / (ms * 1e-3), 4),
            "traffic": traffic, "algorithmic": f"{terms} deviance terms per launch "
            "(sum over references of clipped search box x patch volume)",
            "avg_ms": round(ms, 3)}


def exact_roofline(plan, w, ms, traffic, order: str = "serial") -> dict:
    """Exact-plane variance against the FP32 roof.  `achieved` is the
    algorithmic rate: the multiply-adds of the union-voxel form,
    2*q*P*n*P per group, over the launch's time, whichever order ran.  The
    serial order (the reference's, every multiply and add a separate float32
    VALU operation) executes those and 3*q*P*U more; the pair order executes
    three separable passes per overlapping member pair on the VALU and the
    plane products on the matrix cores, reported as `executed`."""
    sec = ms * 1e-3
    ach = w["flops_mac"] / sec / 1e12
    out = {"kernel": exact_kernel(plan, order), "bound": "valu-fp32", "achieved": round(ach, 3),
           "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(ach / PEAK_FP32_TFLOPS, 4),
           "traffic": traffic, "order": order,
           "algorithmic": "float32 2*q*P*n*P per group (q = min(k, n) planes): {} flops per "
                          "launch ({} entries, {} union voxels, {} overlapping member "
                          "pairs)".format(w["flops_mac"], w["entries"], w["union_voxels"],
                                          w["pairs"]),
           "avg_ms": round(ms, 3)}
    if order == "pairs":
        out["executed"] = {
            "valu_tflops": round(w["flops_pairs"] / sec / 1e12, 3),
            "mfma_tflops": round(w["flops_pairs_mfma"] / sec / 1e12, 3),
            "valu_flops": w["flops_pairs"], "mfma_flops": w["flops_pairs_mfma"],
            "note": "per pair 2*P*(b0+b1+b2) flops of separable passes (VALU) and 2*4*P of "
                    "plane products (v_mfma_f32_4x4x1_16b_f32)"}
    else:
        out["executed"] = {"valu_tflops": round(w["flops"] / sec / 1e12, 3),
                           "valu_flops": w["flops"],
                           "note": "q*P*(2*n*P + 3*U) separately rounded float32 operations"}
    return out


def graph_leg(pipe, vol, steps: int) -> dict:
    """Both stages as one captured HIP graph against call-by-call launches
    (same kernels, bit-identical output), per two-stage call."""
    import torch
    n = max(steps, 20)
    pipe.run_graph(vol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n):
        pipe.run_graph(vol)
    torch.cuda.synchronize()
    g_ms = 1e3 * (time.perf_counter() - t0) / n
    t0 = time.perf_counter()
    for _ in range(n):
        pipe.run(vol)
    torch.cuda.synchronize()
    e_ms = 1e3 * (time.perf_counter() - t0) / n
    same = bool(torch.equal(pipe.run_graph(vol).view(torch.int32),
                            pipe.run(vol)[1].est.view(torch.int32)))
    return {"graph_ms": round(g_ms, 3), "call_by_call_ms": round(e_ms, 3),
            "value": round(vol.numel() / (g_ms * 1e-3), 1), "speedup": round(e_ms / g_ms, 3),
            "bit_identical": same}


def surrogate_leg(profile, shape, dev, peaks=(2.0, 10.0, 50.0), reps: int = 2) -> dict:
    """The learned pilot against the plain stage 1 on the bench phantom
    family at several count levels: per peak the two-stage time of each,
    their ratio, the output PSNRs against the noiseless phantom, and which
    pilot the surrogate's check kept."""
    import torch

    from bmnd_amd import surrogate as sg
    from bmnd_amd.engine import PoissonPipeline
    from bmnd_amd.phantoms import ellipsoid_volume, psnr
    if not os.path.exists(sg.DEFAULT_WEIGHTS):
        return {"skipped": "no trained network"}
    net = sg.PilotNet.load(sg.DEFAULT_WEIGHTS, dev)
    check = sg.PilotCheck()
    pipe = PoissonPipeline(profile, shape, dev)
    res = {}
    for peak in peaks:
        truth, noisy = ellipsoid_volume(shape, seed=0, peak=peak)
        v = torch.from_numpy(noisy).to(dev)
        pipe.run(v)
        pipe.run_surrogate(v, net, check)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            _, wi = pipe.run(v)
        torch.cuda.synchronize()
        plain_ms = 1e3 * (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        for _ in range(reps):
            st, info = pipe.run_surrogate(v, net, check)
        torch.cuda.synchronize()
        surr_ms = 1e3 * (time.perf_counter() - t0) / reps
        p_plain = psnr(wi.est.cpu().numpy(), truth)
        p_surr = psnr(st.wiener.est.cpu().numpy(), truth)
        res[f"{peak:g}"] = {
            "plain_ms": round(plain_ms, 2), "surrogate_ms": round(surr_ms, 2),
            "value": round(truth.size / (surr_ms * 1e-3), 1),
            "ratio": round(plain_ms / surr_ms, 4), "psnr_plain_db": round(p_plain, 3),
            "psnr_surrogate_db": round(p_surr, 3), "psnr_loss_db": round(p_plain - p_surr, 3),
            "pilot": info.used,
            # the agreement check runs only where the count gate lets the network in
            "check_distance": round(info.distance, 4) if np.isfinite(info.distance) else None,
            "mean_count": round(info.mean_count, 4)}
    return {"network": os.path.basename(sg.DEFAULT_WEIGHTS), "params": net.num_parameters,
            "peaks": res, "threshold": check.threshold,
            "max_mean_count": check.max_mean_count}


class _Prefixed:
    """A KernelTimer view that prefixes span names (one per batch)."""

    def __init__(self, timer, prefix):
        self.timer, self.prefix = timer, prefix

    def start(self, name):
        return self.timer.start(f"{self.prefix}.{name}")

    def stop(self, tok):
        self.timer.stop(tok)


def run_batch(args) -> int:
    """Batched workloads: one BatchPipeline per dimensionality, a step runs
    every batch once."""
    import torch

    from bmnd_amd import BM3DProfile, BM4DProfile, NoiseModel, phantoms
    from bmnd_amd.engine import (BatchPipeline, KernelTimer, blockmatch_work, exact_std_work,
                                 table_first_order)

    dev = torch.device("cuda", 0)
    if args.config == "batch-uneven":
        imgs, itruth, vols, vtruth = phantoms.uneven_batch(0)
        groups = [("2d", BM3DProfile(noise_model=NoiseModel.POISSON), imgs, itruth),
                  ("3d", BM4DProfile(noise_model=NoiseModel.POISSON), vols, vtruth)]
        cfg = {"workload": "batch of 32 images (sides uniform in [96,1024]) and 16 volumes "
                           "(sides uniform in [48,128]), peaks uniform in [1,50], one batch "
                           "per dimensionality",
               "inputs": [list(x.shape) for x in imgs + vols], "seed": 0,
               "profile": "BM3DProfile / BM4DProfile(noise_model=POISSON)"}
    else:
        profile, noisy_list, truth_list, cfg = make_inputs(args.config)
        groups = [("batch", profile, noisy_list, truth_list)]
        cfg = dict(cfg, batched=True)
    runs = []
    for tag, prof, inputs, truths in groups:
        order = table_first_order(inputs)  # as bmnd_batch orders a batch
        inputs = [inputs[i] for i in order]
        truths = [truths[i] for i in order] if truths else truths
        bp = BatchPipeline(prof, [x.shape for x in inputs], dev, deviance=args.deviance)
        runs.append((tag, bp, bp.pack(inputs), inputs, truths))
    voxels = sum(bp.total_voxels for _, bp, _, _, _ in runs)
    last = {}

    def step(timer=None):
        for tag, bp, packed, _, _ in runs:
            ht, wi = bp.run(packed, _Prefixed(timer, tag) if timer else None)
            last[tag] = (ht, wi)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    timer = KernelTimer()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(timer)
    torch.cuda.synchronize()
    ms_per_step = 1e3 * (time.perf_counter() - t0) / args.steps
    per_kernel = {k: float(np.mean(v)) for k, v in timer.totals_ms().items()}

    rooflines, psnr = {}, {}
    for tag, bp, _, inputs, truths in runs:
        ht, wi = last[tag]
        for st, out in (("ht", ht), ("wiener", wi)):
            stage = bp.stages[st]
            terms = sum(blockmatch_work(stage.plan, x.shape, sc)[1]
                        for x, sc in zip(inputs, stage.schedules))
            ms = per_kernel[f"{tag}.{st}.blockmatch"]
            ach = terms / (ms * 1e-3) / 1e12
            rooflines[f"{tag}.{st}.blockmatch"] = {
                "bound": "transcendental", "achieved": round(ach, 4),
                "peak": round(TERM_RATE / 1e12, 4), "unit": "Tterm/s",
                "frac": round(ach * 1e12 / TERM_RATE, 4), "traffic": None,
                "algorithmic": f"{terms} deviance terms", "avg_ms": round(ms, 3)}
            if f"{tag}.{st}.exact_std" in per_kernel:
                flops = 0
                for i, x in enumerate(inputs):
                    r0, r1 = stage.ref_off[i], stage.ref_off[i + 1]
                    flops += exact_std_work(stage.plan, bp.shape3s[i], out.sel[r0:r1],
                                            out.counts[r0:r1], bp._single.k)["flops_mac"]
                ms = per_kernel[f"{tag}.{st}.exact_std"]
                ach = flops / (ms * 1e-3) / 1e12
                rooflines[f"{tag}.{st}.exact_std"] = {
                    "bound": "valu-fp32", "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(ach / PEAK_FP32_TFLOPS, 4), "traffic": None,
                    "algorithmic": f"float32 2*q*P*n*P: {flops} flops", "avg_ms": round(ms, 3)}
        if truths:
            from bmnd_amd.phantoms import psnr as _psnr
            outs = bp.unpack(wi.est)
            psnr[tag] = {"noisy_db": round(float(np.mean([_psnr(x, t) for x, t in
                                                         zip(inputs, truths)])), 3),
                         "out_db": round(float(np.mean([_psnr(o, t) for o, t in
                                                       zip(outs, truths)])), 3)}
    dominant = max(per_kernel, key=lambda k: per_kernel[k])
    roof = rooflines.get(dominant) or {"bound": None, "achieved": None, "peak": None,
                                       "unit": None, "frac": None, "traffic": None}
    roof = dict(roof, kernel=dominant)
    cpu = None
    if not args.no_cpu_baseline:
        first = runs[0][3][0]
        cpu = cpu_baseline(args.config, first, args.cpu_sample)
        sample = cpu.pop("_sample")
        cpu["gpu_on_sample"] = gpu_on_sample(runs[0][1]._single.profile, sample)
        cpu["gpu_over_cpu_on_sample"] = round(cpu["gpu_on_sample"]["value"] / cpu["value"], 1)
    value = voxels / (ms_per_step * 1e-3)
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "voxels/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32", "data": "synthetic",
        "config": dict(cfg, pilot="plain", deviance=args.deviance, voxels=voxels),
        "roofline": roof, "rooflines": rooflines, "cpu_baseline": cpu,
        "kernels_ms": {k: round(v, 3) for k, v in sorted(per_kernel.items())},
        "psnr": psnr,
    }
    line["graph"] = batch_graph_leg(runs, last, args.steps)
    print(json.dumps(line), flush=True)
    return 0


def batch_graph_leg(runs, last, steps: int) -> dict:
    """Every batch of the step captured as ONE HIP graph (engine.BatchGraph)
    against the call-by-call step (same kernels, bit-identical outputs)."""
    import torch

    from bmnd_amd.engine import BatchGraph
    g = BatchGraph([bp for _, bp, _, _, _ in runs])
    packed = [p for _, _, p, _, _ in runs]
    n = max(steps, 5)
    g.run(packed)  # capture
    g.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n):
        g.run()
    torch.cuda.synchronize()
    g_ms = 1e3 * (time.perf_counter() - t0) / n
    same = all(torch.equal(g.outputs[j].view(torch.int32), last[tag][1].est.view(torch.int32))
               and torch.equal(g.pilots[j].view(torch.int32), last[tag][0].est.view(torch.int32))
               for j, (tag, _, _, _, _) in enumerate(runs))
    voxels = sum(bp.total_voxels for _, bp, _, _, _ in runs)
    out = {"graph_ms": round(g_ms, 3), "value": round(voxels / (g_ms * 1e-3), 1),
           "stage1_modes": [m[0] for m in g.modes], "bit_identical": same}
    g.release()
    return out




# ---- the stationary Gaussian model --------------------------------------------
GAUSS_SIGMA = 0.1


def gaussian_inputs(config: str):
    from bmnd_amd import BM3DProfile, BM4DProfile, phantoms
    rng = np.random.default_rng(7)
    if config == "gauss-bm4d-256":
        truths = [phantoms.ellipsoid_volume((256, 256, 256), seed=0, peak=10.0)[0]]
        prof, desc = BM4DProfile(), "256^3 BM4D Gaussian, ellipsoid phantom in [0, 1], sigma 0.1"
    else:
        truths = [phantoms.blob_image(512, seed=s)[0] for s in range(16)]
        prof, desc = BM3DProfile(), "16 x 512^2 BM3D Gaussian, blob phantoms in [0, 1], sigma 0.1"
    truths = [(t / float(t.max())).astype(np.float32) for t in truths]
    noisy = [(t + GAUSS_SIGMA * rng.standard_normal(t.shape)).astype(np.float32) for t in truths]
    return prof, noisy, truths, {"workload": desc, "shape": list(noisy[0].shape) if len(noisy) == 1
                                 else [len(noisy)] + list(noisy[0].shape),
                                 "profile": type(prof).__name__ + "()", "sigma": GAUSS_SIGMA}


def gaussian_oracle_params(profile, shape, model, threads):
    """The oracle's parameters for the Gaussian profile with the product's
    noise description (its host tables)."""
    import oracle
    from bmnd_amd.engine import make_plan
    from bmnd_amd.gaussianvariance import blockmatch_noise_ssd, cross_covariance_maps
    from bmnd_amd.profilevalidation import resolve_profile
    from bmnd_amd.transforms import spatial_matrices
    res = resolve_profile(profile, shape)
    p = oracle.default_params(len(shape), shape, threads)
    maps, plans = {}, {}
    for st in ("ht", "wiener"):
        plans[st] = make_plan(profile, st, len(shape))
        mats = [spatial_matrices(plans[st].spatial, b)[0] for b in plans[st].block]
        maps[st] = cross_covariance_maps(model, plans[st].block, mats)
    gamma = 0.0 if res.conventional_nf else float(profile.gamma)
    noise = (blockmatch_noise_ssd(model, plans["ht"].block, plans["ht"].radius3[3 - len(shape):])
             if gamma else None)
    oracle.set_gaussian(p, global_sigma=model.global_sigma, domain=model.variance_domain_shape,
                        ht_maps=maps["ht"], wi_maps=maps["wiener"],
                        ht_match=float(profile.ht_match_threshold),
                        wi_match=float(profile.wiener_match_threshold), ht_noise_ssd=noise,
                        gamma=gamma, k=res.variance_k)
    return p


def gaussian_variance_work(plan, counts, k: int) -> int:
    """Algorithmic float32 operations of a stage's exact-plane variance
    launch: per group of n members and plane q < min(k, n) with m nonzero
    row entries, P * (3 m + 4 m (m - 1) / 2) (a diagonal term is two
    multiplies and an add, a pair term an add, two multiplies and an add)."""
    from bmnd_amd.transforms import group_matrices
    hist = np.bincount(counts.cpu().numpy().astype(np.int64), minlength=plan.max_m + 1)
    total = 0
    for n in range(1, plan.max_m + 1):
        if not hist[n]:
            continue
        G = group_matrices(n)[0]
        per = 0
        for q in range(min(k, n)):
            m = int(np.count_nonzero(G[q]))
            per += 3 * m + 2 * m * (m - 1)
        total += int(hist[n]) * per * plan.P
    return total


def run_gaussian(args) -> int:
    import torch

    import oracle
    from bmnd_amd.engine import KernelTimer, blockmatch_work
    from bmnd_amd.gaussian import GaussianPipeline
    from bmnd_amd.gaussianvariance import build_variance_model
    from bmnd_amd.phantoms import psnr as _psnr
    from bmnd_amd.profilevalidation import resolve_profile
    profile, noisy_list, truth_list, cfg = gaussian_inputs(args.config)
    dev = torch.device("cuda", 0)
    vols = [torch.from_numpy(v).to(dev) for v in noisy_list]
    voxels = sum(int(v.numel()) for v in vols)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pipe = GaussianPipeline(profile, noisy_list[0].shape, dev)
    pipe.run(vols[0], sigma=GAUSS_SIGMA)
    torch.cuda.synchronize()
    cold_ms = 1e3 * (time.perf_counter() - t0)
    last = {}

    def step(timer=None):
        outs = []
        for v in vols:
            ht, wi, _ = pipe.run(v, sigma=GAUSS_SIGMA, timer=timer)
            last["ht"], last["wiener"] = ht, wi
            outs.append(wi.est)
        return outs

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    timer = KernelTimer()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        outs = step(timer)
    torch.cuda.synchronize()
    ms_per_step = 1e3 * (time.perf_counter() - t0) / args.steps
    per_kernel = {k: float(np.mean(v)) for k, v in timer.totals_ms().items()}
    rooflines = {}
    for st in ("ht", "wiener"):
        plan = pipe.plans[st]
        pairs, terms = blockmatch_work(plan, noisy_list[0].shape)
        ms = per_kernel[f"{st}.blockmatch"]
        flops = 3 * terms + (2 * pairs if (st == "ht" and pipe.gamma) else 0)
        ach = flops / (ms * 1e-3) / 1e12
        rooflines[f"{st}.blockmatch"] = {
            "kernel": "k_blockmatch<{}, {}, {}, {}, {}, {}, 10, 512>(BMArgs)".format(
                *plan.block3, *plan.radius3),
            "bound": "valu-fp32", "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": round(ach / PEAK_FP32_TFLOPS, 4), "traffic": None,
            "algorithmic": f"float32 SSD: 3 operations per term, {terms} terms ({pairs} "
                           "candidates) per launch" + (", 2 per candidate for the noise-bias "
                                                       "correction" if st == "ht" and pipe.gamma
                                                       else ""),
            "avg_ms": round(ms, 3)}
        if f"{st}.gaussian_variance" in per_kernel:
            ms = per_kernel[f"{st}.gaussian_variance"]
            w = gaussian_variance_work(plan, last[st].counts, pipe.k)
            ach = w / (ms * 1e-3) / 1e12
            rooflines[f"{st}.gaussian_variance"] = {
                "kernel": "k_gaussian_var<{}, {}, {}>(GaussVarArgs)".format(
                    plan.P, EXACT_CONFIG[tuple(plan.block3)], 128 if plan.P > 64 else 64),
                "bound": "valu-fp32", "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": round(ach / PEAK_FP32_TFLOPS, 4), "traffic": None,
                "algorithmic": f"{w} float32 operations per launch (pair-lag terms of the "
                               f"first {pipe.k} planes)", "avg_ms": round(ms, 3)}
    dominant = max((k for k in per_kernel if k != "gaussian.tables"), key=lambda k: per_kernel[k])
    roof = rooflines.get(dominant) or {"kernel": dominant, "bound": None, "achieved": None,
                                      "peak": None, "unit": None, "frac": None, "traffic": None,
                                      "avg_ms": round(per_kernel[dominant], 3)}
    out0 = outs[0].cpu().numpy()
    psnr = {"noisy_db": round(_psnr(noisy_list[0], truth_list[0], 1.0), 3),
            "out_db": round(_psnr(out0, truth_list[0], 1.0), 3)}
    cpu = None
    if not args.no_cpu_baseline:
        threads = host_threads()
        side = min(args.cpu_sample or (64 if noisy_list[0].ndim == 3 else 256), *noisy_list[0].shape)
        sample = crop(noisy_list[0], side)
        res = resolve_profile(profile, sample.shape)
        model = build_variance_model(sample, sigma=GAUSS_SIGMA, nf=res.effective_nf)
        params = gaussian_oracle_params(profile, sample.shape, model, threads)
        oracle.build()
        t0 = time.perf_counter()
        o_out, _ = oracle.bmnd_gaussian(sample, params)
        dt = time.perf_counter() - t0
        spipe = GaussianPipeline(profile, sample.shape, dev)
        sv = torch.from_numpy(sample).to(dev)
        spipe.run(sv, sigma=GAUSS_SIGMA)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(3):
            g_out = spipe.run(sv, sigma=GAUSS_SIGMA)[1].est
        torch.cuda.synchronize()
        g_dt = (time.perf_counter() - t1) / 3
        g_np = g_out.cpu().numpy()
        cpu = {"value": float(sample.size / dt), "unit": "voxels/s", "cores": threads,
               "kind": "port", "seconds": round(dt, 3),
               "sample": f"{side}^{sample.ndim} centre crop of the first input, oracle "
                         "Gaussian pipeline (both stages)",
               "gpu_on_sample": {"value": float(sample.size / g_dt), "ms": round(1e3 * g_dt, 3)},
               "gpu_over_cpu_on_sample": round(dt / g_dt, 1),
               "gpu_vs_oracle_bits_differing": int(np.sum(g_np.view(np.int32)
                                                          != o_out.view(np.int32)))}
    line = {
        "metric": "voxels/s of the Gaussian bmnd() on 1 MI355X",
        "value": round(voxels / (ms_per_step * 1e-3), 1), "unit": "voxels/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic", "config": dict(cfg, k=pipe.k, gamma=pipe.gamma,
                                            nf=list(pipe.nf) if pipe.nf else None),
        "roofline": roof, "rooflines": rooflines, "cpu_baseline": cpu,
        "cold_call_ms": round(cold_ms, 1),
        "kernels_ms": {k: round(v, 3) for k, v in sorted(per_kernel.items())},
        "psnr": psnr}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
