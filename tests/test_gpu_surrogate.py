"""The pilot network's device kernels against the float64 reference
(oracle/pilot_net.py, pinned by finite differences on the CPU), on the same
bf16-rounded inputs and weights:
- each MFMA convolution configuration (bias + ReLU, bf16 out): within the
  bf16 rounding of its output (2^-8 relative);
- the float-output (backward-data) configurations and the flipped weights;
- the weight gradients: within float32 accumulation error (1e-4 relative);
- the autotuner's tile variants: bit-identical outputs;
- the whole network's output and every parameter gradient: within the error
  of bf16 activations (stated per test);
- training on the device lowers the loss;
- the surrogate pilot falls back to the plain stage 1 (bit-identical to
  bmnd()) when its network disagrees with the plain pilot of sub-volumes,
  and, with the packaged trained network, stays within 0.1 dB PSNR of the
  plain pipeline."""

import ctypes
import os

import numpy as np
import pytest
import torch

from bmnd_amd import BM4DProfile, NoiseModel, _lib, bmnd, phantoms
from bmnd_amd import surrogate as sg
from oracle import pilot_net as pn

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def bf16_round(a: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def to_bf16_tensor(a: np.ndarray) -> torch.Tensor:
    bits = (np.ascontiguousarray(bf16_round(a)).view(np.uint32) >> 16).astype(np.uint16)
    return torch.from_numpy(bits.view(np.int16)).to(DEV)


def from_bf16_tensor(t: torch.Tensor) -> np.ndarray:
    bits = t.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    return bits.view(np.float32)


def pack_weights(w: np.ndarray) -> np.ndarray:
    co, _, cin = w.shape
    out = np.zeros((co, sg.kpad(cin)), np.float32)
    out[:, :27 * cin] = w.reshape(co, -1)
    return out


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream(DEV).cuda_stream)


P = sg._ptr

FWD = [(16, 0, 16, False), (16, 0, 32, False), (32, 0, 32, False), (32, 0, 64, False),
       (64, 0, 64, False), (64, 32, 32, True), (32, 16, 16, True)]
FLOAT = [(16, 16), (32, 16), (32, 32), (64, 32), (64, 64), (32, 96), (16, 48)]
GEOM = (2, 4, 6, 32)


def _inputs(ca, cb, up, seed):
    rng = np.random.default_rng(seed)
    B, D, H, W = GEOM
    xa = bf16_round(rng.normal(0.5, 1.0, ((B, D // 2, H // 2, W // 2) if up else (B, D, H, W))
                               + (ca,)))
    xb = bf16_round(rng.normal(0.5, 1.0, (B, D, H, W, cb))) if cb else None
    xin = pn.up2(xa.astype(np.float64)) if up else xa.astype(np.float64)
    if cb:
        xin = np.concatenate([xin, xb.astype(np.float64)], axis=-1)
    return xa, xb, xin


def _forward(xa, xb, ca, cb, up, w, bias, co, out_float, mt):
    B, D, H, W = GEOM
    out = torch.empty((B, D, H, W, co), dtype=torch.float32 if out_float else torch.int16,
                      device=DEV)
    used = ctypes.c_int32(-1)
    # device copies held until the kernel has run
    ta, tw = to_bf16_tensor(xa), to_bf16_tensor(pack_weights(w))
    tb = None if xb is None else to_bf16_tensor(xb)
    tbias = None if bias is None else torch.from_numpy(bias).to(DEV)
    _lib.check(_lib.lib().bmnd_conv3d_forward(
        P(ta), P(tb), ca, cb, int(up), P(tw), P(tbias), co, P(out), int(out_float), B, D, H, W,
        mt, ctypes.byref(used), stream()))
    torch.cuda.synchronize()
    res = out.cpu().numpy() if out_float else from_bf16_tensor(out)
    return res, used.value


@pytest.mark.parametrize("cfg", FWD, ids=[f"{a}+{b}->{c}{'u' if u else ''}" for a, b, c, u in FWD])
def test_conv_forward_matches_float64(cfg):
    ca, cb, co, up = cfg
    print(f"conv forward {cfg}", flush=True)
    rng = np.random.default_rng(10 + ca + cb + co)
    xa, xb, xin = _inputs(ca, cb, up, 1 + co)
    w = bf16_round(rng.normal(0, np.sqrt(2 / (27 * (ca + cb))), (co, 27, ca + cb)))
    b = rng.normal(0, 0.1, co).astype(np.float32)
    ref = pn.relu(pn.conv3d(xin, w.astype(np.float64), b.astype(np.float64)))
    got, _ = _forward(xa, xb, ca, cb, up, w, b, co, False, 0)
    err = np.abs(got - ref)
    assert np.all(err <= 2.0 ** -8 * np.abs(ref) + 1e-5 * np.abs(ref).max()), float(err.max())


@pytest.mark.parametrize("cfg", FLOAT, ids=[f"{a}->{c}" for a, c in FLOAT])
def test_conv_float_output_matches_float64(cfg):
    ca, co = cfg
    print(f"conv float out {cfg}", flush=True)
    rng = np.random.default_rng(20 + ca + co)
    xa, _, xin = _inputs(ca, 0, False, 3 + co)
    w = bf16_round(rng.normal(0, np.sqrt(2 / (27 * ca)), (co, 27, ca)))
    ref = pn.conv3d(xin, w.astype(np.float64), None)
    got, _ = _forward(xa, None, ca, 0, False, w, None, co, True, 0)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.abs(ref).max()


def test_autotuned_tiles_are_bit_identical():
    print("conv tiles", flush=True)
    rng = np.random.default_rng(5)
    xa, xb, _ = _inputs(64, 32, True, 7)
    w = bf16_round(rng.normal(0, 0.05, (32, 27, 96)))
    b = rng.normal(0, 0.1, 32).astype(np.float32)
    outs = [_forward(xa, xb, 64, 32, True, w, b, 32, False, mt)[0] for mt in (1, 2)]
    tuned, used = _forward(xa, xb, 64, 32, True, w, b, 32, False, 0)
    assert used in (1, 2)  # W = 32 allows tiles of 16 and 32 voxels
    for o in outs:
        assert np.array_equal(o.view(np.int32), tuned.view(np.int32))


def test_flipped_weights_give_the_input_gradient():
    print("conv flip", flush=True)
    rng = np.random.default_rng(6)
    ci, co = 48, 16
    B, D, H, W = GEOM
    x = rng.normal(size=(B, D, H, W, ci))
    w = bf16_round(rng.normal(0, 0.05, (co, 27, ci)))
    gy = bf16_round(rng.normal(size=(B, D, H, W, co)))
    ref = pn.conv3d_backward(x, w.astype(np.float64), gy.astype(np.float64))[0]
    wt = torch.empty(ci * sg.kpad(co), dtype=torch.int16, device=DEV)
    tw, tg = to_bf16_tensor(pack_weights(w)), to_bf16_tensor(gy)
    _lib.check(_lib.lib().bmnd_conv_flip_weights(P(tw), ci, co, P(wt), stream()))
    out = torch.empty((B, D, H, W, ci), dtype=torch.float32, device=DEV)
    _lib.check(_lib.lib().bmnd_conv3d_forward(P(tg), None, co, 0, 0, P(wt), None,
                                              ci, P(out), 1, B, D, H, W, 0, None, stream()))
    got = out.cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.abs(ref).max()


WG = [(16, 0, 16, False), (16, 0, 32, False), (32, 0, 32, False), (32, 0, 64, False),
      (64, 0, 64, False), (64, 32, 32, True), (32, 16, 16, True)]


@pytest.mark.parametrize("cfg", WG, ids=[f"{a}+{b}->{c}{'u' if u else ''}" for a, b, c, u in WG])
def test_weight_gradient_matches_float64(cfg):
    ca, cb, co, up = cfg
    print(f"wgrad {cfg}", flush=True)
    rng = np.random.default_rng(30 + ca + co)
    xa, xb, xin = _inputs(ca, cb, up, 9 + co)
    B, D, H, W = GEOM
    gy = bf16_round(rng.normal(size=(B, D, H, W, co)))
    ref = pn.conv3d_backward(xin, np.zeros((co, 27, ca + cb)), gy.astype(np.float64))[1]
    ref = pack_weights(ref)
    L = _lib.lib()
    chunk = 128
    nparts = L.bmnd_conv3d_wgrad_parts(B, D, H, W, chunk)
    parts = torch.empty(nparts * co * sg.kpad(ca + cb), dtype=torch.float32, device=DEV)
    grad = torch.empty((co, sg.kpad(ca + cb)), dtype=torch.float32, device=DEV)
    ta, tg = to_bf16_tensor(xa), to_bf16_tensor(gy)
    tb = None if xb is None else to_bf16_tensor(xb)
    _lib.check(L.bmnd_conv3d_wgrad(P(ta), P(tb), ca, cb, int(up), P(tg), co, P(parts), chunk,
                                   P(grad), B, D, H, W, stream()))
    got = grad.cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 1e-4 * np.abs(ref).max()
    assert np.all(got[:, 27 * (ca + cb):] == 0)


def _net_from(params):
    net = sg.PilotNet(DEV, seed=0)
    for name, *_ in pn.LAYERS:
        w = params[name + ".w"]
        net.set_layer(name, w, params[name + ".b"])
    # the device computes with bf16 weights in the MFMA layers: use those
    used = {}
    for name, *_ in pn.LAYERS:
        w, b = net.layer_weights(name)
        used[name + ".w"], used[name + ".b"] = w.astype(np.float64), b.astype(np.float64)
    return net, used


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


def test_network_output_and_gradients_match_float64():
    print("network gradients", flush=True)
    rng = np.random.default_rng(11)
    net, params = _net_from(pn.init_params(12, scale=1.0))
    x = rng.normal(2.5, 0.8, size=(1, 8, 8, 64)).astype(np.float32)
    t = (x + rng.normal(0, 0.3, size=x.shape)).astype(np.float32)
    xt = torch.from_numpy(x).to(DEV)
    y, acts = net.forward(xt, keep=True)
    sums, gy, inside = net.loss_and_grad(y, torch.from_numpy(t).to(DEV), margin=1)
    net.backward(gy, acts)
    torch.cuda.synchronize()
    y_ref, a_ref = pn.forward(params, x.astype(np.float64))
    l_ref, gy_ref = pn.loss(y_ref, t.astype(np.float64), 1)
    g_ref = pn.backward(params, a_ref, gy_ref)
    yo = y.cpu().numpy()
    # bf16 activations at every layer: relative error of the residual branch
    assert _rel(yo - x, y_ref - x) < 2e-2
    assert float(sums.sum().item()) / inside == pytest.approx(l_ref, rel=5e-2)
    worst = {}
    for name, *_ in pn.LAYERS:
        for part in (".w", ".b"):
            got = net.view(name + part, net.grad).cpu().numpy()
            ref = g_ref[name + part]
            if part == ".w" and name not in sg.SMALL:
                got = got[:, :ref.size // ref.shape[0]].reshape(ref.shape)
            worst[name + part] = _rel(got, ref)
    print("gradient relative errors:", {k: round(v, 4) for k, v in worst.items()})
    assert max(worst.values()) < 0.08, worst


def test_training_on_the_device_lowers_the_loss():
    print("training", flush=True)
    rng = np.random.default_rng(13)
    truth = np.zeros((8, 16, 16, 64), np.float32)
    zz, yy, xx = np.meshgrid(np.arange(16), np.arange(16), np.arange(64), indexing="ij")
    for i in range(8):
        truth[i] = 4.0 + 6.0 * (((zz - 8) ** 2 + (yy - 8) ** 2 + (xx - 8 * i) ** 2) < 40)
    noisy = rng.poisson(truth).astype(np.float32)
    X = sg.anscombe(torch.from_numpy(noisy).to(DEV))
    T = sg.anscombe(torch.from_numpy(truth).to(DEV))
    net = sg.PilotNet(DEV, seed=2)
    from bmnd_amd.surrogate_train import train
    hist = train(net, X, T, epochs=70, batch=2, lr=4e-3, margin=2, val_frac=0.25, log_loss=True,
                 log=lambda m: print(m, flush=True))
    assert hist["val_end"] < 0.8 * hist["val_start"], hist


@pytest.mark.parametrize("peak", [1.5, 10.0])
def test_surrogate_falls_back_to_the_plain_pilot(peak):
    """A network whose pilot is far off is caught by the sub-volume check at
    a low count level; at a high one the network is not run.  Either way the
    output is bmnd()'s, bit for bit."""
    print(f"surrogate fallback at peak {peak}", flush=True)
    _, noisy = phantoms.ellipsoid_volume((48, 52, 64), seed=21, peak=peak, n_ellipsoids=8,
                                         axes_range=(4.0, 14.0))
    prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    net = sg.PilotNet(DEV, seed=0)
    net.set_layer("c1e", np.zeros((16, 27)), np.array([3.0]))  # pilot far off
    out, info = sg.bmnd_surrogate(noisy, prof, net, check=sg.PilotCheck(side=40, count=3),
                                  return_info=True)
    assert info.used == "plain"
    if info.mean_count <= sg.PilotCheck.max_mean_count:
        assert info.distance > info.threshold
    else:
        assert np.isnan(info.distance)
    assert (info.mean_count <= 0.25) == (peak < 5)
    ref = bmnd(noisy, BM4DProfile(noise_model=NoiseModel.POISSON))
    assert np.array_equal(out.view(np.int32), ref.view(np.int32))


@pytest.mark.skipif(not os.path.exists(sg.DEFAULT_WEIGHTS), reason="no trained network packaged")
@pytest.mark.parametrize("peak", [2.0, 20.0])
def test_trained_network_meets_the_bar_or_falls_back(peak):
    """The packaged network keeps its pilot where the final output stays
    within 0.1 dB of the plain pipeline (it is better at 2 counts) and falls
    back to the plain stage 1 where it would not (20 counts)."""
    print(f"trained surrogate at peak {peak}", flush=True)
    truth, noisy = phantoms.ellipsoid_volume((112, 112, 128), seed=7, peak=peak,
                                             n_ellipsoids=14, axes_range=(4.0, 28.0))
    prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    out, info = sg.bmnd_surrogate(noisy, prof, return_info=True)
    plain = bmnd(noisy, BM4DProfile(noise_model=NoiseModel.POISSON))
    gap = phantoms.psnr(plain, truth) - phantoms.psnr(out, truth)
    print(f"trained surrogate: used {info.used}, check {info.distance:.3f}, mean count "
          f"{info.mean_count:.3f}, gap {gap:.3f} dB")
    if peak <= 2.0:
        assert info.used == "network"
        assert gap <= 0.1
    else:
        assert info.used == "plain"
        assert np.array_equal(out.view(np.int32), plain.view(np.int32))
