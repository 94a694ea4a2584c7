"""numpy restatement of the learned coarse-grid correction network: forward
pass, loss and its gradient with respect to every weight.

TEST INFRASTRUCTURE ONLY (what the network's HIP kernels are checked
against); the product never imports it.

Network: y = sum_t sum_ci W3[t][ci] h2[ci](x + o_t),
h2 = relu(sum_t sum_ci W2[t][ci][co] h1[ci](x + o_t)),
h1 = relu(sum_t (W1[t][0][c] p(x + o_t) + W1[t][1][c] (r p)(x + o_t))),
taps o_t = (dx, dz) with t = (dx + 1) * 3 + (dz + 1), the fields zero
outside their array, r = 1 / vel_t.  Loss: mean over the region of
(p + y - target)^2.

With emulate=True the BF16 roundings of the device are reproduced: every
product reads its operands rounded to BF16 (round to nearest even) -- the
inputs p and r p (r p formed in FP32), W1, h1, W2, h2 and W3 -- and in the
backward pass dz2 is rounded to BF16 before the two products that read it;
everything else is FP64.  With emulate=False the network is plain FP64 (for
finite-difference checks)."""
import numpy as np

C = 16


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), as float64."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


NW1 = 9 * 2 * C
NPARAMS = NW1 + 9 * C * C + 9 * C


def split(theta):
    th = np.asarray(theta, np.float64)
    w1 = th[:NW1].reshape(9, 2, C)
    w2 = th[NW1:NW1 + 9 * C * C].reshape(9, C, C)
    w3 = th[NW1 + 9 * C * C:].reshape(9, C)
    return w1, w2, w3


def shift(a, dx, dz):
    """out[..., i, j] = a[..., i + dx, j + dz], zero outside."""
    out = np.zeros_like(a)
    nx, nz = a.shape[-2:]
    xs, xd = (slice(dx, nx), slice(0, nx - dx)) if dx >= 0 else (slice(0, nx + dx), slice(-dx, nx))
    zs, zd = (slice(dz, nz), slice(0, nz - dz)) if dz >= 0 else (slice(0, nz + dz), slice(-dz, nz))
    out[..., xd, zd] = a[..., xs, zs]
    return out


TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def forward(p, r, theta, emulate=True):
    """p, r: [nx][nz] field and 1 / vel_t.  Returns y and the intermediates."""
    w1, w2, w3 = split(theta)
    p = np.asarray(p, np.float64)
    q = np.float64(np.float32(np.asarray(r, np.float32) * np.asarray(p, np.float32)))  # r p formed in FP32
    rnd = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    pin, qin, w1e, w2e, w3e = rnd(p), rnd(q), rnd(w1), rnd(w2), rnd(w3)
    sp = np.stack([np.stack([shift(pin, dx, dz), shift(qin, dx, dz)]) for dx, dz in TAPS])  # [t][k][x][z]
    z1 = np.einsum("tkc,tkxz->cxz", w1e, sp)
    h1 = rnd(np.maximum(z1, 0))
    s1 = np.stack([shift(h1, dx, dz) for dx, dz in TAPS])          # [t][ci][x][z]
    z2 = np.einsum("tio,tixz->oxz", w2e, s1)
    h2 = rnd(np.maximum(z2, 0))
    s2 = np.stack([shift(h2, dx, dz) for dx, dz in TAPS])
    y = np.einsum("ti,tixz->xz", w3e, s2)
    return y, dict(sp=sp, h1=h1, s1=s1, z2=z2, h2=h2, s2=s2, w2e=w2e, w3e=w3e)


def loss_grad(p, rv, target, theta, region, emulate=True):
    """Mean over the region of res^2, res = p + y - target, summed over a
    stack p[e][nx][nz] with models rv[e]; returns (loss, gradient)."""
    x0, x1, z0, z1 = region
    p = np.asarray(p, np.float64)
    target = np.asarray(target, np.float64)
    E = p.shape[0]
    count = (x1 - x0) * (z1 - z0) * E
    w1, w2, w3 = split(theta)
    g1 = np.zeros_like(w1)
    g2 = np.zeros_like(w2)
    g3 = np.zeros_like(w3)
    loss = 0.0
    rv = np.asarray(rv, np.float32)
    for e in range(E):
        y, c = forward(p[e], rv[e], theta, emulate)
        r = p[e] + y - target[e]
        mask = np.zeros_like(r, bool)
        mask[x0:x1, z0:z1] = True
        r = np.where(mask, r, 0.0)
        loss += float((r * r).sum())
        dy = 2 * r / count
        # layer 3
        g3 += np.einsum("xz,tixz->ti", dy, c["s2"])
        dh2 = sum(c["w3e"][t][:, None, None] * shift(dy, -dx, -dz)[None] for t, (dx, dz) in enumerate(TAPS))
        dz2 = np.where(c["h2"] > 0, dh2, 0.0)
        if emulate:
            dz2 = bf16(dz2)
        # layer 2
        g2 += np.einsum("tixz,oxz->tio", c["s1"], dz2)
        dh1 = sum(np.einsum("io,oxz->ixz", c["w2e"][t], shift(dz2, -dx, -dz)) for t, (dx, dz) in enumerate(TAPS))
        dz1 = np.where(c["h1"] > 0, dh1, 0.0)
        # layer 1
        g1 += np.einsum("tkxz,cxz->tkc", c["sp"], dz1)
    grad = np.concatenate([g1.ravel(), g2.ravel(), g3.ravel()])
    return loss / count, grad
