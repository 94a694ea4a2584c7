"""Pins the network oracle (oracle/convnet.py) on the CPU: its gradient
against central finite differences of its own FP64 loss, its BF16 rounding
against the IEEE definition, and the convolution against a direct loop."""
import numpy as np

from oracle import convnet as cn


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.14159, 65504.0, 1e-30], np.float32)
    r = cn.bf16(x)
    # 1 + 2^-8 is half-way between 1 and 1 + 2^-7: ties to even -> 1
    assert r[0] == 1.0 and r[1] == 1.0
    assert r[2] == np.float32(1.0078125)  # above the half-way point
    assert r[3] == np.float32(1.0078125)
    for v, w in zip(x, r):  # at most half an ulp of an 8-bit significand: 2^-8 relative
        assert abs(float(w) - float(v)) <= 2.0 ** -8 * abs(float(v))


def test_forward_matches_direct_loop():
    rng = np.random.default_rng(0)
    theta = rng.standard_normal(cn.NPARAMS) * 0.3
    p = rng.standard_normal((9, 8)).astype(np.float32)
    rv = rng.uniform(0.05, 0.3, (9, 8)).astype(np.float32)
    y, _ = cn.forward(p, rv, theta, emulate=False)
    w1, w2, w3 = cn.split(theta)
    nx, nz = p.shape
    q = (rv * p).astype(np.float64)

    def at(a, i, j):
        return a[i, j] if 0 <= i < nx and 0 <= j < nz else 0.0

    h1 = np.zeros((16, nx, nz))
    for i in range(nx):
        for j in range(nz):
            for c in range(16):
                h1[c, i, j] = max(0.0, sum(w1[t, 0, c] * at(p, i + dx, j + dz) + w1[t, 1, c] * at(q, i + dx, j + dz)
                                           for t, (dx, dz) in enumerate(cn.TAPS)))
    h2 = np.zeros((16, nx, nz))
    for i in range(nx):
        for j in range(nz):
            for o in range(16):
                s = 0.0
                for t, (dx, dz) in enumerate(cn.TAPS):
                    for ci in range(16):
                        s += w2[t, ci, o] * at(h1[ci], i + dx, j + dz)
                h2[o, i, j] = max(0.0, s)
    want = np.zeros((nx, nz))
    for i in range(nx):
        for j in range(nz):
            want[i, j] = sum(w3[t, ci] * at(h2[ci], i + dx, j + dz) for t, (dx, dz) in enumerate(cn.TAPS)
                             for ci in range(16))
    assert np.allclose(y, want, rtol=1e-12, atol=1e-12)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    theta = rng.standard_normal(cn.NPARAMS) * 0.2
    p = rng.standard_normal((2, 14, 12))
    rv = rng.uniform(0.05, 0.3, p.shape).astype(np.float32)
    t = p + 0.1 * rng.standard_normal(p.shape)
    region = (5, 9, 5, 8)
    loss, g = cn.loss_grad(p, rv, t, theta, region, emulate=False)
    # every W1 and W3 weight, a spread of W2 weights
    n1, n2 = cn.NW1, 9 * 16 * 16
    idx = list(range(0, n1)) + list(range(n1, n1 + n2, 37)) + list(range(n1 + n2, cn.NPARAMS))
    h = 1e-6

    def L(th):
        return cn.loss_grad(p, rv, t, th, region, emulate=False)[0]

    for k in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (L(tp) - L(tm)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])
    assert loss > 0
