"""Pins of the float64 pilot-network reference (oracle/pilot_net.py) on the
CPU: its hand-written gradients against central finite differences of its
own loss (every layer, weights and biases), its convolution against a
direct triple loop, and pooling/upsampling adjointness.  The device kernels
are then checked against this reference (tests/test_gpu_surrogate.py)."""

import numpy as np
import pytest

from oracle import pilot_net as pn


def test_conv3d_against_direct_loops():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(1, 3, 4, 5, 2))
    w = rng.normal(size=(3, 27, 2))
    b = rng.normal(size=3)
    y = pn.conv3d(x, w, b)
    ref = np.zeros_like(y)
    for z in range(3):
        for yy in range(4):
            for xx in range(5):
                for co in range(3):
                    s = b[co]
                    for t, (dz, dy, dx) in enumerate(pn.TAPS):
                        zz, y2, x2 = z + dz, yy + dy, xx + dx
                        if 0 <= zz < 3 and 0 <= y2 < 4 and 0 <= x2 < 5:
                            s += w[co, t] @ x[0, zz, y2, x2]
                    ref[0, z, yy, xx, co] = s
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_pool_and_upsample_adjoints():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 4, 6, 8, 3))
    g = rng.normal(size=(2, 2, 3, 4, 3))
    # <pool(x), g> = <x, pool^T(g)> and <up(g), x> = <g, up^T(x)>
    assert np.isclose(np.sum(pn.pool2(x) * g), np.sum(x * pn.pool2_backward(g)))
    assert np.isclose(np.sum(pn.up2(g) * x), np.sum(g * pn.up2_backward(x)))


def test_network_gradients_against_finite_differences():
    rng = np.random.default_rng(2)
    params = pn.init_params(3)
    x = rng.normal(2.0, 1.0, size=(1, 4, 4, 4))
    t = x + rng.normal(0.0, 0.3, size=x.shape)

    def f(p):
        y, _ = pn.forward(p, x)
        return pn.loss(y, t, 0)[0]

    y, acts = pn.forward(params, x)
    _, gy = pn.loss(y, t, 0)
    grads = pn.backward(params, acts, gy)
    h = 1e-6
    for key in params:
        flat = params[key].reshape(-1)
        picks = rng.choice(flat.size, size=min(4, flat.size), replace=False)
        for j in picks:
            old = flat[j]
            flat[j] = old + h
            fp = f(params)
            flat[j] = old - h
            fm = f(params)
            flat[j] = old
            num = (fp - fm) / (2 * h)
            ana = grads[key].reshape(-1)[j]
            assert abs(num - ana) <= 1e-6 + 1e-4 * abs(num), (key, j, num, ana)


def test_loss_margin_excludes_the_faces():
    y = np.zeros((1, 8, 8, 8))
    t = np.ones_like(y)
    l0, g0 = pn.loss(y, t, 0)
    l2, g2 = pn.loss(y, t, 2)
    assert l0 == pytest.approx(1.0) and l2 == pytest.approx(1.0)
    assert np.count_nonzero(g2) == 4 ** 3 and np.count_nonzero(g0) == 8 ** 3
