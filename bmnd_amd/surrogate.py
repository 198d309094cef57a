"""Learned stage-1 pilot for 3-D Poisson volumes (no counterpart in the
reference).

The reference's stage 2 takes three things from stage 1 (bmndalgo.py:636-644
the Wiener variance model, :660-667 the pilot patches for the gains, :714-721
the volume it matches on), all from the hard-threshold pilot `step1_est`.
`bmnd_surrogate` supplies that pilot from a small convolutional network
instead and runs the reference's stage 2 on it unchanged.

Network: a 3-level U-Net in the Anscombe domain A(x) = 2 sqrt(x + 3/8)
(unit noise variance at every count level), standardised per volume
(u = (A(x) - mean) / std, which keeps the bf16 activations near unit
scale), residual output:
    (A(pilot) - mean) / std ~ u + f(u),
f = c1e . c1d . c1c[up(d2) | s1] . ... with 3x3x3 "same" convolutions and
ReLU, channels 16 / 32 / 64 at full, half and quarter resolution (2x2x2 mean
pooling down, nearest 2x upsampling up, skip concatenation), 353,697
parameters.  Every layer runs in the library's hand-written kernels
(include/bmnd_conv.h): bf16 MFMA convolutions with float32 accumulation for
the 16-to-96-channel layers, VALU for the 1-channel first and last layers,
float32 master weights and Adam.

Fallback: before the pilot is used, the plain stage 1 runs on a few
sub-volumes of the input (one batched launch per kernel) and the network's
pilot must agree with theirs, away from the sub-volume borders, to within a
set RMS distance in the Anscombe domain; otherwise the plain stage 1 runs on
the whole input and its pilot is used.
"""

from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib

# (name, input channels from the first source, from the skip, output
# channels, upsampled first source, pyramid level)
LAYERS = (
    ("c1a", 1, 0, 16, False, 0),
    ("c1b", 16, 0, 16, False, 0),
    ("c2a", 16, 0, 32, False, 1),
    ("c2b", 32, 0, 32, False, 1),
    ("c3a", 32, 0, 64, False, 2),
    ("c3b", 64, 0, 64, False, 2),
    ("c2c", 64, 32, 32, True, 1),
    ("c2d", 32, 0, 32, False, 1),
    ("c1c", 32, 16, 16, True, 0),
    ("c1d", 16, 0, 16, False, 0),
    ("c1e", 16, 0, 1, False, 0),
)
SMALL = ("c1a", "c1e")  # 1-channel layers: float32 [16][27] weights on VALU
TAPS = 27
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                               "pilot_net_bm4d.npz")


def kpad(cin: int) -> int:
    return (TAPS * cin + 31) // 32 * 32


def _ptr(t: torch.Tensor | None) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _align(n: int, a: int = 64) -> int:
    return (n + a - 1) // a * a


@dataclass
class _Param:
    offset: int      # first element in the flat parameter buffer
    shape: tuple     # logical shape of the stored block
    size: int


def anscombe(x: torch.Tensor, inverse: bool = False, center: float = 0.0,
             spread: float = 1.0) -> torch.Tensor:
    """(A(x) - center) / spread, A(x) = 2 sqrt(x + 3/8), or its inverse."""
    out = torch.empty_like(x)
    _lib.check(_lib.lib().bmnd_anscombe(_ptr(x), _ptr(out), x.numel(), int(inverse), center,
                                        spread, _stream(x.device)))
    return out


def anscombe_stats(x: torch.Tensor, with_max: bool = False):
    """Mean and standard deviation of A(x) over a count volume (float64
    sums on the device, read back): the standardisation the network's
    input and output use; with_max, also the volume's maximum."""
    part = torch.empty(3072, dtype=torch.float64, device=x.device)
    sums = torch.empty(3, dtype=torch.float64, device=x.device)
    _lib.check(_lib.lib().bmnd_anscombe_moments(_ptr(x), x.numel(), _ptr(part), _ptr(sums),
                                                _stream(x.device)))
    s1, s2, mx = (float(v) for v in sums.cpu().tolist())
    n = float(x.numel())
    mean = s1 / n
    std = max(max((s2 / n - mean * mean), 0.0) ** 0.5, 1e-3)
    return (mean, std, mx) if with_max else (mean, std)


def _stream(device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class PilotNet:
    """The pilot network on one device.  Parameters live in one flat float32
    buffer (each block 64-element aligned) with a bf16 copy the MFMA layers
    read; a layer's weights are stored [COUT][KP] (k = tap * CIN + ci), the
    1-channel layers' as [16][27], biases after each layer's weights."""

    def __init__(self, device, seed: int = 0):
        self.device = torch.device(device)
        self.params: dict[str, _Param] = {}
        off = 0
        for name, ca, cb, co, _, _ in LAYERS:
            if name in SMALL:
                wshape = (16, TAPS)
            else:
                wshape = (co, kpad(ca + cb))
            wsize = int(np.prod(wshape))
            self.params[name + ".w"] = _Param(off, wshape, wsize)
            off = _align(off + wsize)
            self.params[name + ".b"] = _Param(off, (co,), co)
            off = _align(off + co)
        self.numel = off
        self.master = torch.zeros(off, dtype=torch.float32, device=self.device)
        self.bf16 = torch.zeros(off, dtype=torch.int16, device=self.device)
        self.grad = torch.zeros(off, dtype=torch.float32, device=self.device)
        self._m = self._v = None
        self.steps = 0
        self._flipped: dict[str, torch.Tensor] = {}
        self.init(seed)

    # ---- parameters ---------------------------------------------------------
    @property
    def num_parameters(self) -> int:
        """Trainable parameters (weights without padding columns, biases)."""
        n = 0
        for name, ca, cb, co, _, _ in LAYERS:
            n += 16 * TAPS + co if name in SMALL else co * TAPS * (ca + cb) + co
        return n

    def view(self, key: str, buf: torch.Tensor | None = None) -> torch.Tensor:
        p = self.params[key]
        b = self.master if buf is None else buf
        return b[p.offset:p.offset + p.size].view(p.shape)

    def init(self, seed: int) -> None:
        """He-normal weights (fan-in 27 * CIN), zero biases, and a zero last
        layer so that the untrained network is the identity pilot."""
        rng = np.random.default_rng(seed)
        host = np.zeros(self.numel, dtype=np.float32)
        for name, ca, cb, co, _, _ in LAYERS:
            p = self.params[name + ".w"]
            cin = ca + cb
            if name == "c1e":
                w = np.zeros(p.shape, np.float32)
            elif name in SMALL:
                w = rng.normal(0.0, np.sqrt(2.0 / TAPS), p.shape).astype(np.float32)
            else:
                w = np.zeros(p.shape, np.float32)
                w[:, :TAPS * cin] = rng.normal(0.0, np.sqrt(2.0 / (TAPS * cin)),
                                               (co, TAPS * cin))
            host[p.offset:p.offset + p.size] = w.reshape(-1)
        self.load_flat(host)

    def load_flat(self, host: np.ndarray) -> None:
        self.master.copy_(torch.from_numpy(np.ascontiguousarray(host, np.float32)))
        self._sync_bf16()

    def _sync_bf16(self) -> None:
        _lib.check(_lib.lib().bmnd_to_bf16(_ptr(self.master), _ptr(self.bf16), self.numel,
                                           _stream(self.device)))
        self._flipped.clear()

    def save(self, path: str) -> None:
        arrays = {k: self.view(k).cpu().numpy() for k in self.params}
        np.savez(path, steps=np.int64(self.steps), **arrays)

    @classmethod
    def load(cls, path: str = DEFAULT_WEIGHTS, device=None) -> "PilotNet":
        dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
        net = cls(dev)
        data = np.load(path, allow_pickle=False)
        host = np.zeros(net.numel, dtype=np.float32)
        for k, p in net.params.items():
            a = np.asarray(data[k], dtype=np.float32)
            if a.shape != p.shape:
                raise ValueError(f"{path}: {k} has shape {a.shape}, expected {p.shape}")
            host[p.offset:p.offset + p.size] = a.reshape(-1)
        net.steps = int(data["steps"])
        net.load_flat(host)
        return net

    def set_layer(self, name: str, weight: np.ndarray, bias: np.ndarray) -> None:
        """Weights in the logical layout [COUT][27][CIN] ([16][27] for the
        1-channel layers) and the bias, e.g. from a float64 reference."""
        host = self.master.cpu().numpy().copy()
        p = self.params[name + ".w"]
        if name in SMALL:
            w = np.asarray(weight, np.float32).reshape(p.shape)
        else:
            co, kp = p.shape
            w = np.zeros(p.shape, np.float32)
            w[:, :weight.size // co] = np.asarray(weight, np.float32).reshape(co, -1)
        host[p.offset:p.offset + p.size] = w.reshape(-1)
        pb = self.params[name + ".b"]
        host[pb.offset:pb.offset + pb.size] = np.asarray(bias, np.float32)
        self.load_flat(host)

    def layer_weights(self, name: str) -> tuple[np.ndarray, np.ndarray]:
        """(weights [COUT][27][CIN] or [16][27], bias) as float32 on the host,
        rounded to bf16 where the layer computes in bf16."""
        lay = {x[0]: x for x in LAYERS}[name]
        _, ca, cb, co, _, _ = lay
        if name in SMALL:
            w = self.view(name + ".w").cpu().numpy().copy()
        else:
            raw = self.view(name + ".w", self.bf16).cpu().numpy().astype(np.uint16)
            w = (raw.astype(np.uint32) << 16).view(np.float32)[:, :TAPS * (ca + cb)]
            w = w.reshape(co, TAPS, ca + cb)
        return w, self.view(name + ".b").cpu().numpy().copy()

    def _w16(self, name: str) -> torch.Tensor:
        return self.view(name + ".w", self.bf16)

    def _flip(self, name: str, cin: int, co: int) -> torch.Tensor:
        if name not in self._flipped:
            wt = torch.empty(cin * kpad(co), dtype=torch.int16, device=self.device)
            _lib.check(_lib.lib().bmnd_conv_flip_weights(_ptr(self._w16(name)), cin, co, _ptr(wt),
                                                         _stream(self.device)))
            self._flipped[name] = wt
        return self._flipped[name]

    # ---- layers ---------------------------------------------------------------
    def _act(self, B, D, H, W, C):
        return torch.empty((B, D, H, W, C), dtype=torch.int16, device=self.device)

    def _conv(self, name, a, b, shape, mt=0):
        lay = {x[0]: x for x in LAYERS}[name]
        _, ca, cb, co, up, _ = lay
        B, D, H, W = shape
        out = self._act(B, D, H, W, co)
        _lib.check(_lib.lib().bmnd_conv3d_forward(
            _ptr(a), _ptr(b), ca, cb, int(up), _ptr(self._w16(name)),
            _ptr(self.view(name + ".b")), co, _ptr(out), 0, B, D, H, W, mt, None,
            _stream(self.device)))
        return out

    def _pool(self, x):
        B, D, H, W, C = x.shape
        out = self._act(B, D // 2, H // 2, W // 2, C)
        _lib.check(_lib.lib().bmnd_conv_pool2(_ptr(x), _ptr(out), C, B, D, H, W,
                                              _stream(self.device)))
        return out

    def forward(self, x: torch.Tensor, keep: bool = False, mt: int = 0):
        """x: float32 [B, D, H, W] in the Anscombe domain (D, H multiples of 4,
        W of 64).  Returns y [B, D, H, W] and, with keep, the activations."""
        if x.dtype != torch.float32 or x.dim() != 4 or not x.is_contiguous():
            raise ValueError("x must be a contiguous float32 [B, D, H, W] tensor")
        B, D, H, W = (int(s) for s in x.shape)
        if D % 4 or H % 4 or W % 64:
            raise ValueError("the pilot network needs D and H divisible by 4 and W by 64")
        L = _lib.lib()
        s = _stream(self.device)
        g1, g2, g3 = (B, D, H, W), (B, D // 2, H // 2, W // 2), (B, D // 4, H // 4, W // 4)
        a1 = self._act(*g1, 16)
        _lib.check(L.bmnd_conv_in_forward(_ptr(x), _ptr(self.view("c1a.w")),
                                          _ptr(self.view("c1a.b")), _ptr(a1), *g1, s))
        s1 = self._conv("c1b", a1, None, g1, mt)
        p1 = self._pool(s1)
        a2 = self._conv("c2a", p1, None, g2, mt)
        s2 = self._conv("c2b", a2, None, g2, mt)
        p2 = self._pool(s2)
        a3 = self._conv("c3a", p2, None, g3, mt)
        b3 = self._conv("c3b", a3, None, g3, mt)
        u2 = self._conv("c2c", b3, s2, g2, mt)
        d2 = self._conv("c2d", u2, None, g2, mt)
        u1 = self._conv("c1c", d2, s1, g1, mt)
        d1 = self._conv("c1d", u1, None, g1, mt)
        y = torch.empty_like(x)
        _lib.check(L.bmnd_conv_out_forward(_ptr(d1), _ptr(self.view("c1e.w")),
                                           _ptr(self.view("c1e.b")), _ptr(x), _ptr(y), *g1, s))
        if keep:
            acts = dict(x=x, a1=a1, s1=s1, p1=p1, a2=a2, s2=s2, p2=p2, a3=a3, b3=b3, u2=u2,
                        d2=d2, u1=u1, d1=d1)
            return y, acts
        return y

    # ---- backward -------------------------------------------------------------
    def backward(self, gy: torch.Tensor, acts: dict, chunk: int = 4096) -> None:
        """Gradients of every parameter into self.grad from dL/dy (float32
        [B, D, H, W]) and the activations forward(keep=True) returned."""
        L = _lib.lib()
        s = _stream(self.device)
        dev = self.device
        B, D, H, W = (int(v) for v in gy.shape)
        g1, g2, g3 = (B, D, H, W), (B, D // 2, H // 2, W // 2), (B, D // 4, H // 4, W // 4)
        vox = {0: B * D * H * W, 1: B * (D // 2) * (H // 2) * (W // 2),
               2: B * (D // 4) * (H // 4) * (W // 4)}
        geo = {0: g1, 1: g2, 2: g3}
        self.grad.zero_()
        need = -(-vox[0] // chunk) * 16 * TAPS  # the 1-channel layers
        for _, ca, cb, co, _, lvl in LAYERS:
            if co > 1 and ca > 1:
                need = max(need, -(-vox[lvl] // chunk) * co * kpad(ca + cb), -(-vox[lvl] // chunk) * co)
        parts = torch.empty(need, dtype=torch.float32, device=dev)

        def gw(name):
            return self.view(name + ".w", self.grad)

        def gb(name):
            return self.view(name + ".b", self.grad)

        def bias_grad(name, g_bf16, lvl, co):
            _lib.check(L.bmnd_conv_bias_grad(_ptr(g_bf16), None, co, _ptr(parts), chunk,
                                             _ptr(gb(name)), vox[lvl], s))

        def wgrad(name, xa, xb, gpre):
            lay = {x[0]: x for x in LAYERS}[name]
            _, ca, cb, co, up, lvl = lay
            _lib.check(L.bmnd_conv3d_wgrad(_ptr(xa), _ptr(xb), ca, cb, int(up), _ptr(gpre), co,
                                           _ptr(parts), chunk, _ptr(gw(name)), *geo[lvl], s))
            bias_grad(name, gpre, lvl, co)

        def back_data(name, gpre):
            """float32 gradient of the layer's (virtual) input."""
            lay = {x[0]: x for x in LAYERS}[name]
            _, ca, cb, co, _, lvl = lay
            cin = ca + cb
            out = torch.empty(geo[lvl] + (cin,), dtype=torch.float32, device=dev)
            _lib.check(L.bmnd_conv3d_forward(_ptr(gpre), None, co, 0, 0, _ptr(self._flip(name, cin, co)),
                                             None, cin, _ptr(out), 1, *geo[lvl], 0, None, s))
            return out

        def mask(g, act):
            out = torch.empty(act.shape, dtype=torch.int16, device=dev)
            _lib.check(L.bmnd_conv_mask_grad(_ptr(g), _ptr(act), _ptr(out), act.numel(), s))
            return out

        a = acts
        # last layer
        _lib.check(L.bmnd_conv_small_wgrad(1, None, _ptr(a["d1"]), _ptr(gy), None, _ptr(parts),
                                           chunk, _ptr(gw("c1e")), *g1, s))
        _lib.check(L.bmnd_conv_bias_grad(None, _ptr(gy), 1, _ptr(parts), chunk, _ptr(gb("c1e")),
                                         vox[0], s))
        g_d1 = torch.empty(a["d1"].shape, dtype=torch.int16, device=dev)
        _lib.check(L.bmnd_conv_out_backward(_ptr(gy), _ptr(self.view("c1e.w")), _ptr(a["d1"]),
                                            _ptr(g_d1), *g1, s))
        # level 1 decoder
        wgrad("c1d", a["u1"], None, g_d1)
        g_u1 = mask(back_data("c1d", g_d1), a["u1"])
        wgrad("c1c", a["d2"], a["s1"], g_u1)
        gx_c1c = back_data("c1c", g_u1)                   # [.., 48]: up(d2) | s1
        g_d2 = torch.empty(a["d2"].shape, dtype=torch.int16, device=dev)
        _lib.check(L.bmnd_conv_up_grad(_ptr(gx_c1c), 48, _ptr(a["d2"]), _ptr(g_d2), 32, *g1, s))
        # level 2 decoder
        wgrad("c2d", a["u2"], None, g_d2)
        g_u2 = mask(back_data("c2d", g_d2), a["u2"])
        wgrad("c2c", a["b3"], a["s2"], g_u2)
        gx_c2c = back_data("c2c", g_u2)                   # [.., 96]: up(b3) | s2
        g_b3 = torch.empty(a["b3"].shape, dtype=torch.int16, device=dev)
        _lib.check(L.bmnd_conv_up_grad(_ptr(gx_c2c), 96, _ptr(a["b3"]), _ptr(g_b3), 64, *g2, s))
        # level 3
        wgrad("c3b", a["a3"], None, g_b3)
        g_a3 = mask(back_data("c3b", g_b3), a["a3"])
        wgrad("c3a", a["p2"], None, g_a3)
        g_p2 = back_data("c3a", g_a3)                     # float [.., 32] at level 3
        # level 2 encoder: s2 feeds the pool and the c2c skip
        g_s2 = torch.empty(a["s2"].shape, dtype=torch.int16, device=dev)
        _lib.check(L.bmnd_conv_skip_grad(_ptr(gx_c2c), 96, 64, _ptr(g_p2), _ptr(a["s2"]),
                                         _ptr(g_s2), 32, *g2, s))
        wgrad("c2b", a["a2"], None, g_s2)
        g_a2 = mask(back_data("c2b", g_s2), a["a2"])
        wgrad("c2a", a["p1"], None, g_a2)
        g_p1 = back_data("c2a", g_a2)                     # float [.., 16] at level 2
        # level 1 encoder: s1 feeds the pool and the c1c skip
        g_s1 = torch.empty(a["s1"].shape, dtype=torch.int16, device=dev)
        _lib.check(L.bmnd_conv_skip_grad(_ptr(gx_c1c), 48, 32, _ptr(g_p1), _ptr(a["s1"]),
                                         _ptr(g_s1), 16, *g1, s))
        wgrad("c1b", a["a1"], None, g_s1)
        g_a1 = mask(back_data("c1b", g_s1), a["a1"])
        _lib.check(L.bmnd_conv_small_wgrad(0, _ptr(a["x"]), None, None, _ptr(g_a1), _ptr(parts),
                                           chunk, _ptr(gw("c1a")), *g1, s))
        bias_grad("c1a", g_a1, 0, 16)

    def loss_and_grad(self, y: torch.Tensor, target: torch.Tensor, margin: int = 4,
                      log_loss: bool = False, eps: float = 1e-4):
        """(per-sample sums of squared errors over the interior [B] (device),
        dL/dy, interior voxels per sample).  L is the batch mean squared
        error, or with log_loss the sum over samples of log(mse + eps)."""
        B, D, H, W = (int(v) for v in y.shape)
        gy = torch.empty_like(y)
        parts = torch.empty(B * (-(-(D * H * W) // 256)), dtype=torch.float32, device=self.device)
        sums = torch.empty(B, dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().bmnd_conv_mse(_ptr(y), _ptr(target), _ptr(gy), _ptr(parts), margin,
                                            B, D, H, W, int(log_loss), eps, _ptr(sums),
                                            _stream(self.device)))
        inside = (D - 2 * margin) * (H - 2 * margin) * (W - 2 * margin)
        return sums, gy, inside

    def adam(self, lr: float, beta1: float = 0.9, beta2: float = 0.999,
             eps: float = 1e-8) -> None:
        if self._m is None:
            self._m = torch.zeros_like(self.master)
            self._v = torch.zeros_like(self.master)
        self.steps += 1
        _lib.check(_lib.lib().bmnd_adam_step(_ptr(self.master), _ptr(self.grad), _ptr(self._m),
                                             _ptr(self._v), _ptr(self.bf16), self.numel, lr,
                                             beta1, beta2, eps, self.steps,
                                             _stream(self.device)))
        self._flipped.clear()

    # ---- inference ------------------------------------------------------------
    def pilot(self, vol: torch.Tensor) -> torch.Tensor:
        """The network's stage-1 pilot of a float32 [D, H, W] count volume:
        edge-padded to the network's multiples, Anscombe transform
        standardised by the volume's own mean and spread, forward, inverse,
        cropped back."""
        D, H, W = (int(s) for s in vol.shape)
        Dp, Hp, Wp = -(-D // 4) * 4, -(-H // 4) * 4, -(-W // 64) * 64
        L = _lib.lib()
        s = _stream(self.device)
        x = vol
        if (Dp, Hp, Wp) != (D, H, W):
            x = torch.empty((Dp, Hp, Wp), dtype=torch.float32, device=self.device)
            _lib.check(L.bmnd_pad_edge(_ptr(vol), _ptr(x), D, H, W, Dp, Hp, Wp, 0, s))
        center, spread = anscombe_stats(vol)
        y = self.forward(anscombe(x, center=center, spread=spread).view(1, Dp, Hp, Wp))
        est = anscombe(y.view(Dp, Hp, Wp), inverse=True, center=center, spread=spread)
        if (Dp, Hp, Wp) != (D, H, W):
            out = torch.empty((D, H, W), dtype=torch.float32, device=self.device)
            _lib.check(L.bmnd_pad_edge(_ptr(est), _ptr(out), D, H, W, Dp, Hp, Wp, 1, s))
            est = out
        return est.contiguous()


@dataclass
class PilotCheck:
    """Agreement test of a network pilot with the plain stage 1 on `count`
    sub-volumes of side `side` spread over the input (corners and centre),
    compared `margin` voxels in from their faces (where the plain pilot of
    a sub-volume equals that of the whole input up to the matching window's
    reach).  The distance is the RMS difference of the two pilots in the
    Anscombe domain (where Poisson noise has unit variance at every count
    level), in the worst sub-volume.  The network pilot is kept when that
    distance is at most `threshold` and the input's mean count at most
    `max_mean_count`: the count range where the network was shown to keep
    the final output within 0.1 dB of the plain pipeline."""
    side: int = 48
    count: int = 4
    margin: int = 12
    # calibrated with the packaged network on 256^3 ellipsoid volumes (seed
    # 0) at peaks 2, 5, 10, 20 and 50 counts: distances 0.048, 0.063, 0.075,
    # 0.103, 0.165 against final-output PSNR differences (plain - network)
    # of -0.81, +0.51, +0.68, +1.11, +1.91 dB (mean counts 0.159, 0.398,
    # 0.796, 1.59, 3.98); the distance rises with the count level but varies
    # with the structure (0.059 on a 112^2 x 128 volume at peak 2, mean count
    # 0.106), so the count level is gated directly
    threshold: float = 0.07
    max_mean_count: float = 0.25

    def accepts(self, distance: float, mean_count: float) -> bool:
        return distance <= self.threshold and mean_count <= self.max_mean_count

    _pipes: dict = field(default_factory=dict)

    def origins(self, shape) -> np.ndarray:
        D, H, W = shape
        s = min(self.side, D, H, W)
        spots = [(0, 0, 0), (D - s, H - s, W - s), ((D - s) // 2, (H - s) // 2, (W - s) // 2),
                 (0, H - s, (W - s) // 2), (D - s, 0, 0), ((D - s) // 2, 0, W - s)]
        return np.array(spots[:self.count], dtype=np.int32)

    def distance(self, profile, vol: torch.Tensor, pilot: torch.Tensor) -> float:
        from .engine import BatchPipeline
        shape = tuple(int(x) for x in vol.shape)
        s = min(self.side, *shape)
        org = self.origins(shape)
        key = (repr(profile), shape, s)
        if key not in self._pipes:
            self._pipes[key] = BatchPipeline(profile, [(s, s, s)] * len(org), vol.device)
        bp = self._pipes[key]
        crops = torch.stack([vol[o[0]:o[0] + s, o[1]:o[1] + s, o[2]:o[2] + s] for o in org])
        packed = crops.contiguous().view(-1)
        ht, _, _ = bp.run_ht(packed)
        dev = vol.device
        org_d = torch.from_numpy(org.reshape(-1)).to(dev)
        nb = 16
        parts = torch.empty(len(org) * nb * 2, dtype=torch.float32, device=dev)
        _lib.check(_lib.lib().bmnd_compare_crops(
            _ptr(pilot), _ptr(ht.est), _ptr(org_d), len(org), *shape, s, s, s, self.margin,
            _ptr(parts), nb, _stream(dev)))
        inner = (s - 2 * self.margin) ** 3
        host = parts.view(len(org), nb, 2).cpu().numpy().astype(np.float64)
        return float(np.sqrt(host[:, :, 0].sum(axis=1) / inner).max())


@dataclass
class SurrogateInfo:
    used: str                 # "network" or "plain" (fallback)
    distance: float           # agreement RMS of the check
    threshold: float
    mean_count: float = 0.0   # the input's mean count


def bmnd_surrogate(volume, profile, network: PilotNet | None = None, *,
                   check: PilotCheck | None = None, stage_callback=None,
                   return_info: bool = False):
    """bmnd() with the network's stage-1 pilot (3-D Poisson volumes): the
    reference's argument checks, the network pilot with the plain-pilot
    agreement check (and the plain stage 1 as fallback), then the
    reference's stage 2 on that pilot.  Returns the float32 estimate (and a
    SurrogateInfo with return_info)."""
    from .bmndalgo import _device, _pipeline, _validated_input
    from .engine import check_supported
    vol, resolved = _validated_input(volume, profile, None, None, "auto")
    if vol.ndim != 3:
        raise NotImplementedError("the pilot network is trained for 3-D volumes")
    if not np.any(vol != 0.0):
        return (vol, SurrogateInfo("none", 0.0, 0.0, 0.0)) if return_info else vol
    check_supported(profile, resolved)
    vol = np.ascontiguousarray(vol)
    device = _device()
    net = network if network is not None else default_network(device)
    pipe = _pipeline(profile, vol.shape, device)
    vol_t = torch.from_numpy(vol).to(device)
    on_pilot = None
    if stage_callback is not None:
        def on_pilot(pilot):  # before stage 2, as the reference orders it (bmndalgo.py:632-633)
            stage_callback("ht", pilot.cpu().numpy())
    est, info = pipe.run_surrogate(vol_t, net, check or PilotCheck(), on_pilot=on_pilot)
    out = est.wiener.est.cpu().numpy()
    if stage_callback is not None:
        stage_callback("wiener", out.copy())
    return (out, info) if return_info else out


_DEFAULT: dict = {}


def default_network(device) -> PilotNet:
    """The packaged network (trained on 2000 64^3 crops of ellipsoid
    phantoms, tools in bmnd_amd.surrogate_train), loaded once per device."""
    key = str(device)
    if key not in _DEFAULT:
        if not os.path.exists(DEFAULT_WEIGHTS):
            raise FileNotFoundError(f"{DEFAULT_WEIGHTS} is missing: train the pilot network with "
                                    "`python -m bmnd_amd.surrogate_train`")
        _DEFAULT[key] = PilotNet.load(DEFAULT_WEIGHTS, device)
    return _DEFAULT[key]
