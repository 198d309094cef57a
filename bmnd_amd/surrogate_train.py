"""Train the stage-1 pilot network on the device.

Training pairs: 64^3 crops of Poisson-sampled ellipsoid phantoms (the
256^3 phantom family of bmnd_amd.phantoms, peaks spread geometrically
over the range the network serves) and the plain stage-1 pilot of the
whole volume over the same crop (PoissonPipeline.run_ht, the reference's
hard-threshold pass), both in the Anscombe domain standardised by the
whole observation's mean and spread (as the network is used).  The loss is
the per-crop log mean squared error away from the crop faces (the plain
pilot there depends on voxels outside the crop).  Every forward, backward and optimiser step runs
in the library's kernels (PilotNet).

    python -m bmnd_amd.surrogate_train --out bmnd_amd/data/pilot_net_bm4d.npz
"""

from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np
import torch

from . import BM4DProfile, NoiseModel, phantoms
from .engine import PoissonPipeline
from .surrogate import PilotNet, anscombe, anscombe_stats

EVAL_SEEDS = (0,)  # the benchmark volume's phantom seed, never trained on


def make_pairs(device, n_volumes: int = 8, crops: int = 2000, side: int = 64,
               peaks=(1.5, 60.0), seed: int = 1000, shape=(256, 256, 256), log=print):
    """(inputs, targets): float32 [crops, side, side, side] device tensors
    in the Anscombe domain, crops spread evenly over the volumes."""
    prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    rng = np.random.default_rng(seed)
    pipe = None
    xs, ts = [], []
    per = [crops // n_volumes + (1 if i < crops % n_volumes else 0) for i in range(n_volumes)]
    for i in range(n_volumes):
        peak = peaks[0] * (peaks[1] / peaks[0]) ** (i / max(n_volumes - 1, 1))
        vseed = seed + i
        if vseed in EVAL_SEEDS:
            raise ValueError("training seeds must differ from the evaluation seeds")
        _, noisy = phantoms.ellipsoid_volume(shape, seed=vseed, peak=float(peak))
        v = torch.from_numpy(noisy).to(device)
        if pipe is None:
            pipe = PoissonPipeline(prof, noisy.shape, device)
        ht, _, _ = pipe.run_ht(v)
        # standardised by the observation's statistics, as PilotNet.pilot does
        center, spread = anscombe_stats(v)
        xa = anscombe(v, center=center, spread=spread)
        ta = anscombe(ht.est, center=center, spread=spread)
        org = rng.integers(0, [s - side + 1 for s in shape], size=(per[i], 3))
        for o in org:
            sl = tuple(slice(int(o[d]), int(o[d]) + side) for d in range(3))
            xs.append(xa[sl].clone())
            ts.append(ta[sl].clone())
        log(f"pairs: volume {i + 1}/{n_volumes} (peak {peak:.2f}) -> {per[i]} crops")
    return torch.stack(xs), torch.stack(ts)


def train(net: PilotNet, X: torch.Tensor, T: torch.Tensor, *, epochs: float = 2.0,
          batch: int = 4, lr: float = 1e-3, margin: int = 4, val_frac: float = 0.05,
          seed: int = 0, log_loss: bool = False, log=print) -> dict:
    """Adam with a cosine learning-rate decay over `epochs` passes,
    minimising the batch mean squared error or, with log_loss, the sum of
    the samples' log mean squared errors (each crop's error in decibels
    weighs the same, whatever its count level); returns the loss history
    (mean squared training error every 25 steps, validation mean squared
    error at the start and the end)."""
    n = X.shape[0]
    nval = max(int(round(val_frac * n)), 1)
    g = torch.Generator(device="cpu").manual_seed(seed)
    perm = torch.randperm(n, generator=g)
    val, tr = perm[:nval].to(X.device), perm[nval:]
    steps = int(math.ceil(epochs * len(tr) / batch))

    def val_loss():
        tot, cnt = 0.0, 0
        for i in range(0, nval, batch):
            idx = val[i:i + batch]
            y = net.forward(X[idx].contiguous())
            sums, _, inside = net.loss_and_grad(y, T[idx].contiguous(), margin)
            tot += float(sums.sum().item())
            cnt += inside * len(idx)
        return tot / cnt

    hist = {"val_start": val_loss(), "train": [], "steps": steps}
    log(f"train: {len(tr)} crops, {nval} held out, {steps} steps of {batch}; "
        f"validation loss {hist['val_start']:.5f}")
    order = torch.cat([tr[torch.randperm(len(tr), generator=g)]
                       for _ in range(int(math.ceil(steps * batch / len(tr))) + 1)])
    t0 = time.perf_counter()
    for step in range(steps):
        idx = order[step * batch:(step + 1) * batch].to(X.device)
        xb, tb = X[idx].contiguous(), T[idx].contiguous()
        y, acts = net.forward(xb, keep=True)
        sums, gy, inside = net.loss_and_grad(y, tb, margin, log_loss=log_loss)
        net.backward(gy, acts)
        rate = lr * 0.5 * (1.0 + math.cos(math.pi * step / steps))
        net.adam(rate)
        if step % 25 == 0 or step == steps - 1:
            cur = float(sums.sum().item()) / (inside * len(idx))
            hist["train"].append((step, cur))
            log(f"step {step}/{steps}: loss {cur:.5f} lr {rate:.2e} "
                f"({time.perf_counter() - t0:.1f} s)")
    torch.cuda.synchronize()
    hist["seconds"] = time.perf_counter() - t0
    hist["val_end"] = val_loss()
    log(f"train: {hist['seconds']:.1f} s, validation loss {hist['val_start']:.5f} -> "
        f"{hist['val_end']:.5f}")
    return hist


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--out", required=True)
    ap.add_argument("--volumes", type=int, default=8)
    ap.add_argument("--crops", type=int, default=2000)
    ap.add_argument("--epochs", type=float, default=4.0)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--side", type=int, default=256, help="side of the phantom volumes")
    ap.add_argument("--mse", action="store_true",
                    help="minimise the plain batch mean squared error instead of the "
                         "per-crop log mean squared error")
    args = ap.parse_args(argv)
    if not torch.cuda.is_available():
        print("training needs a GPU", file=sys.stderr)
        return 2
    dev = torch.device("cuda", 0)
    t0 = time.perf_counter()
    X, T = make_pairs(dev, args.volumes, args.crops, shape=(args.side,) * 3,
                      log=lambda m: print(m, flush=True))
    gen_s = time.perf_counter() - t0
    net = PilotNet(dev, seed=0)
    hist = train(net, X, T, epochs=args.epochs, batch=args.batch, lr=args.lr,
                 log_loss=not args.mse, log=lambda m: print(m, flush=True))
    net.save(args.out)
    hist["pairs_seconds"] = gen_s
    with open(args.out + ".json", "w") as f:
        json.dump(hist, f)
    print(json.dumps({"out": args.out, "pairs_s": round(gen_s, 1),
                      "train_s": round(hist["seconds"], 1), "val_start": hist["val_start"],
                      "val_end": hist["val_end"]}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
