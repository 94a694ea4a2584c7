"""numpy restatement of the learned coarse-grid correction network: forward
pass, loss and its gradient with respect to every weight.

TEST INFRASTRUCTURE ONLY (the checker of the network's HIP kernels); the
product never imports it.

Network: y = sum_t sum_ci W3[t][ci] h2[ci](x + o_t),
h2 = relu(sum_t sum_ci W2[t][ci][co] h1[ci](x + o_t)),
h1 = relu(sum_t W1[t][c] p(x + o_t)), taps o_t = (dx, dz) with
t = (dx + 1) * 3 + (dz + 1), the field zero outside its array.  Loss: mean
over the region of (p + y - target)^2.

With emulate=True the BF16 roundings of the device are reproduced: h1 and
W2 enter the 16 -> 16 product rounded to BF16 (round to nearest even), and
in the backward pass dz2 is rounded to BF16 before the two products that
read it; everything else is FP64.  With emulate=False the network is plain
FP64 (for finite-difference checks)."""
import numpy as np

C = 16


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), as float64."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def split(theta):
    th = np.asarray(theta, np.float64)
    w1 = th[:9 * C].reshape(9, C)
    w2 = th[9 * C:9 * C + 9 * C * C].reshape(9, C, C)
    w3 = th[9 * C + 9 * C * C:].reshape(9, C)
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


def forward(p, theta, emulate=True):
    """p: [nx][nz] field.  Returns y and the intermediates."""
    w1, w2, w3 = split(theta)
    p = np.asarray(p, np.float64)
    sp = [shift(p, dx, dz) for dx, dz in TAPS]
    z1 = np.einsum("tc,txz->cxz", w1, np.stack(sp))
    h1 = np.maximum(z1, 0)
    if emulate:
        h1 = bf16(h1)
        w2e = bf16(w2)
    else:
        w2e = w2
    s1 = np.stack([shift(h1, dx, dz) for dx, dz in TAPS])          # [t][ci][x][z]
    z2 = np.einsum("tio,tixz->oxz", w2e, s1)
    h2 = np.maximum(z2, 0)
    s2 = np.stack([shift(h2, dx, dz) for dx, dz in TAPS])
    y = np.einsum("ti,tixz->xz", w3, s2)
    return y, dict(sp=sp, h1=h1, s1=s1, z2=z2, h2=h2, s2=s2, w2e=w2e)


def loss_grad(p, target, theta, region, emulate=True):
    """Mean over the region of r^2, r = p + y - target, summed over a stack
    p[e][nx][nz]; returns (loss, gradient [2592])."""
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
    for e in range(E):
        y, c = forward(p[e], theta, emulate)
        r = p[e] + y - target[e]
        mask = np.zeros_like(r, bool)
        mask[x0:x1, z0:z1] = True
        r = np.where(mask, r, 0.0)
        loss += float((r * r).sum())
        dy = 2 * r / count
        # layer 3
        g3 += np.einsum("xz,tixz->ti", dy, c["s2"])
        dh2 = sum(w3[t][:, None, None] * shift(dy, -dx, -dz)[None] for t, (dx, dz) in enumerate(TAPS))
        dz2 = np.where(c["h2"] > 0, dh2, 0.0)
        if emulate:
            dz2 = bf16(dz2)
        # layer 2
        g2 += np.einsum("tixz,oxz->tio", c["s1"], dz2)
        dh1 = sum(np.einsum("io,oxz->ixz", c["w2e"][t], shift(dz2, -dx, -dz)) for t, (dx, dz) in enumerate(TAPS))
        dz1 = np.where(c["h1"] > 0, dh1, 0.0)
        # layer 1
        g1 += np.einsum("txz,cxz->tc", np.stack(c["sp"]), dz1)
    grad = np.concatenate([g1.ravel(), g2.ravel(), g3.ravel()])
    return loss / count, grad
