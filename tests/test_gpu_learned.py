"""The learned coarse-grid correction on the GPU through the C ABI.

* the network's forward pass, loss and weight gradient (BF16 MFMA, FP32
  accumulation) against the numpy oracle with the same BF16 roundings
  (oracle/convnet.py, itself pinned by finite differences);
* training by Adam on the device lowers the loss of a problem whose exact
  answer is a fixed linear filter;
* the 4th-order coarse stencil of the batch (2x dx, 2x dt) on a field tap of
  the fine 10th-order run: taps are the restriction of the fine field;
* (the gain of the trained network over the plain coarse propagation on a
  held-out model is measured by swfwi_amd.learned and bench.py)."""
import numpy as np
import pytest

from oracle import convnet as cn
from swfwi_amd import capi, learned
from swfwi_amd.batch import Batch, Group, Net

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if capi.device_count() < 1:
        pytest.skip("no HIP device")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def smooth_fields(rng, e, nx, nz):
    from scipy.ndimage import gaussian_filter
    f = np.stack([gaussian_filter(rng.standard_normal((nx, nz)), 2.0) for _ in range(e)])
    return (f / np.abs(f).max()).astype(np.float32)


def test_net_forward_and_gradient_against_oracle():
    rng = np.random.default_rng(3)
    e, nx, nz = 2, 70, 53
    region = (9, 61, 7, 46)  # 52 x 39: partial tiles in both directions
    p = smooth_fields(rng, e, nx, nz)
    rv = rng.uniform(0.05, 0.3, p.shape).astype(np.float32)
    tgt = (p + 0.05 * smooth_fields(rng, e, nx, nz)).astype(np.float32)
    theta = Net.initial(seed=5, scale=0.1)
    net = Net(theta)
    y = net.apply(p, rv, region)
    x0, x1, z0, z1 = region
    for k in range(e):
        want, _ = cn.forward(p[k], rv[k], theta, emulate=True)
        got = y[k]
        assert np.all(got[:x0] == 0) and np.all(got[:, :z0] == 0)
        err = rel_l2(got[x0:x1, z0:z1], want[x0:x1, z0:z1])
        print(f"net forward entry {k}: rel-L2 {err:.2e}", flush=True)
        assert err < 1e-4
    loss, g = net.loss_grad(p, rv, tgt, region)
    wl, wg = cn.loss_grad(p, rv, tgt, theta, region, emulate=True)
    print(f"loss {loss:.6e} oracle {wl:.6e}", flush=True)
    assert abs(loss - wl) <= 1e-4 * wl
    n1, n2 = cn.NW1, 9 * 16 * 16
    for name, sl in (("W1", slice(0, n1)), ("W2", slice(n1, n1 + n2)), ("W3", slice(n1 + n2, cn.NPARAMS))):
        err = rel_l2(g[sl], wg[sl])
        print(f"  d loss / d {name}: rel-L2 {err:.2e}", flush=True)
        assert err < 2e-3, name
    net.close()


def test_net_forward_each_layer2_tap_alone():
    """Layer 2 with all but one of its nine taps zeroed, tap by tap: each
    product of the kernel's layer-2 accumulation chain must contribute on its
    own (a product whose sum is lost or read stale shows up as one wrong tap,
    which the all-taps check can blur into a small overall error)."""
    rng = np.random.default_rng(3)
    nx, nz = 70, 53
    region = (9, 61, 7, 46)
    p = smooth_fields(rng, 1, nx, nz)
    rv = rng.uniform(0.05, 0.3, p.shape).astype(np.float32)
    base = Net.initial(seed=5, scale=0.1)
    x0, x1, z0, z1 = region
    net = Net(base)
    for t in range(9):
        theta = base.copy()
        w2 = theta[cn.NW1:cn.NW1 + 9 * 16 * 16].reshape(9, 16, 16)
        w2[np.arange(9) != t] = 0
        net.set_weights(theta)
        got = net.apply(p, rv, region)[0]
        want, _ = cn.forward(p[0], rv[0], theta, emulate=True)
        err = rel_l2(got[x0:x1, z0:z1], want[x0:x1, z0:z1])
        print(f"layer-2 tap {t} alone: rel-L2 {err:.2e}", flush=True)
        assert np.linalg.norm(want[x0:x1, z0:z1]) > 0
        assert err < 1e-4, t
    net.close()


def test_net_training_lowers_loss():
    """Targets p + F p for a fixed 3x3 filter F: the network can represent the
    correction (F p = relu(F p) - relu(-F p)), so Adam on the device must
    drive the loss well below its starting value."""
    from scipy.ndimage import correlate
    rng = np.random.default_rng(4)
    n, nx, nz = 24, 48, 40
    region = (6, 42, 6, 34)
    p = smooth_fields(rng, n, nx, nz)
    rv = np.full(p.shape, 0.2, np.float32)
    F = np.array([[0.0, 0.1, 0.0], [0.1, -0.4, 0.1], [0.0, 0.1, 0.0]])
    tgt = np.stack([p[k] + correlate(p[k].astype(np.float64), F, mode="constant") for k in range(n)]).astype(
        np.float32)
    net = Net(Net.initial(seed=1, scale=0.01))
    net.set_data(p, rv, tgt, region)
    losses = net.train(batch=8, steps=300, lr=3e-3, seed=7)
    print(f"training loss {losses[0]:.3e} -> {losses[-1]:.3e}", flush=True)
    assert np.all(np.isfinite(losses))
    assert losses[-20:].mean() < 0.1 * losses[:5].mean()
    net.close()


def test_fine_taps_and_coarse_step():
    """A fine-run tap equals the fine field restricted to the coarse grid,
    and a coarse step from two taps stays close to the fine field two fine
    steps on (the 4th-order coarse scheme approximates the same wave)."""
    case = learned.small_case()
    fine, coarse = learned.grids(case)
    v = learned.random_models(case, 1, seed=11)[0]
    fb = Batch([Group(**fine)], snap_prec=capi.SNAP_NONE, order=capi.ORDER_10)
    fb.bind_model(0, 0, learned.expand(v, fine))
    steps = [398, 400, 402]
    fb.set_taps(steps, oz=learned.OZ)
    fb.forward(want_data=False)
    taps = fb.taps()[0]
    assert np.abs(taps).max() > 0
    cb = Batch([Group(**coarse)], snap_prec=capi.SNAP_NONE, order=capi.ORDER_4)
    cb.bind_model(0, 0, learned.expand(learned.restrict(v), coarse))
    prev, cur = cb.step_from(taps[0], taps[1], it0=200, nsteps=1)
    region = learned.region(coarse)
    x0, x1, z0, z1 = region
    c = cur.reshape(taps[2].shape)
    err = rel_l2(c[x0:x1, z0:z1], taps[2][x0:x1, z0:z1])
    print(f"one coarse step from fine taps: rel-L2 against the fine field {err:.3f}", flush=True)
    assert err < 0.2
    assert np.array_equal(prev.reshape(taps[1].shape), taps[1])
