"""Python handle of the mixed-size batch with reduced-precision wavefield
checkpoints (swfwi_batch)."""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from .capi import _VP, _f, check, load


@dataclass
class Group:
    """One grid size: physical nz x nx with sponge nb, nt steps of dt,
    Ricker fm sources at (sx[s], sz[s]), receivers at x = gx0 .. gx0+ng-1 on
    depth gz, `nmodels` models each shot by every source."""
    nx: int
    nz: int
    nb: int
    dx: float
    dt: float
    fm: float
    nt: int
    sx: np.ndarray
    sz: np.ndarray
    gx0: int = 0
    ng: int = 0
    gz: int = 0
    nmodels: int = 1
    dscale: float = 1.0
    extra: dict = field(default_factory=dict)

    @property
    def nxpad(self):
        return self.nx + 2 * self.nb + 12

    @property
    def nzpad(self):
        return self.nz + self.nb + 12

    @property
    def shape(self):
        return (self.nxpad, self.nzpad)

    @property
    def ns(self):
        return int(len(self.sx))


class Batch:
    """Entries (group, model, shot) in that nesting order: every array in and
    out of this class lists them group by group, model by model, shot by shot."""

    def __init__(self, groups, snap_prec=1, order=10):
        self.groups = list(groups)
        dims = np.array([[g.nx, g.nz, g.nb, g.nt, g.ns, g.gx0, g.ng or g.nx, g.gz, g.nmodels] for g in self.groups],
                        np.int32)
        params = np.array([[g.dx, g.dt, g.fm, g.dscale] for g in self.groups], np.float32)
        sx = np.concatenate([np.asarray(g.sx, np.int32) for g in self.groups]).astype(np.int32)
        sz = np.concatenate([np.asarray(g.sz, np.int32) for g in self.groups]).astype(np.int32)
        for g in self.groups:
            g.ng = g.ng or g.nx
        h = _VP()
        check(load().swfwi_batch_create(C.byref(h), len(self.groups), np.ascontiguousarray(dims.ravel()),
                                        np.ascontiguousarray(params.ravel()), sx, sz, int(snap_prec),
                                        int(order)))
        self.h = h
        self.order = int(order)
        self.ntaps = 0
        self.entries = [(gi, m, s) for gi, g in enumerate(self.groups) for m in range(g.nmodels) for s in range(g.ns)]

    def close(self):
        if getattr(self, "h", None):
            load().swfwi_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def set_snapshot_precision(self, p):
        check(load().swfwi_batch_set_snapshot_precision(self.h, int(p)))

    def set_graphs(self, on):
        check(load().swfwi_batch_set_graphs(self.h, int(bool(on))))

    def set_tuning(self, on):
        check(load().swfwi_batch_set_tuning(self.h, int(bool(on))))

    def set_tile_x(self, tx):
        """16, 24 or 32 (fixed), or 0 for the size rule and the tuner."""
        check(load().swfwi_batch_set_tile_x(self.h, int(tx)))

    def set_block_order(self, order):
        """0 entry by entry, 1 a model's shots tile by tile, -1 the rule and the tuner."""
        check(load().swfwi_batch_set_block_order(self.h, int(order)))

    def block_order(self):
        t = C.c_int(0)
        check(load().swfwi_batch_block_order(self.h, C.byref(t)))
        return t.value

    def tile_x(self):
        t = C.c_int(0)
        check(load().swfwi_batch_tile_x(self.h, C.byref(t)))
        return t.value

    def bind_model(self, group, model, vel_t):
        g = self.groups[group]
        v = _f(vel_t)
        assert v.shape == g.shape
        check(load().swfwi_batch_bind_model(self.h, group, model, v))

    def _data_len(self):
        return sum(self.groups[gi].nt * self.groups[gi].ng for gi, _, _ in self.entries)

    def split_data(self, flat):
        """[entry] list of [nt][ng] views of a flat data array."""
        out, o = [], 0
        for gi, _, _ in self.entries:
            g = self.groups[gi]
            n = g.nt * g.ng
            out.append(flat[o:o + n].reshape(g.nt, g.ng))
            o += n
        return out

    def split_grid(self, flat, per_model=False):
        out, o = [], 0
        keys = [(gi, m) for gi, g in enumerate(self.groups) for m in range(g.nmodels)] if per_model else \
            [(gi, m) for gi, m, _ in self.entries]
        for gi, _ in keys:
            g = self.groups[gi]
            n = g.nxpad * g.nzpad
            out.append(flat[o:o + n].reshape(g.shape))
            o += n
        return out

    def forward(self, want_data=True):
        if not want_data:
            check(load().swfwi_batch_forward(self.h, None))
            return None
        d = np.zeros(self._data_len(), np.float32)
        check(load().swfwi_batch_forward(self.h, d.ctypes.data_as(C.c_void_p)))
        return d

    def adjoint_source(self, dobs_flat, want=True):
        v = np.zeros(self._data_len(), np.float32) if want else None
        check(load().swfwi_batch_adjoint_source(self.h, _f(dobs_flat),
                                                v.ctypes.data_as(C.c_void_p) if want else None))
        return v

    def backward(self, vsrc_flat=None, want_entries=True, want_models=True):
        ne = sum(self.groups[gi].nxpad * self.groups[gi].nzpad for gi, _, _ in self.entries)
        nm = sum(g.nxpad * g.nzpad * g.nmodels for g in self.groups)
        ge = np.zeros(ne, np.float32) if want_entries else None
        gm = np.zeros(nm, np.float32) if want_models else None
        vs = None if vsrc_flat is None else _f(vsrc_flat)
        check(load().swfwi_batch_backward(self.h, None if vs is None else vs.ctypes.data_as(C.c_void_p),
                                          None if ge is None else ge.ctypes.data_as(C.c_void_p),
                                          None if gm is None else gm.ctypes.data_as(C.c_void_p)))
        return ge, gm

    def set_taps(self, steps, oz=0):
        steps = np.ascontiguousarray(steps, np.int32)
        check(load().swfwi_batch_set_taps(self.h, int(steps.size), steps, int(oz)))
        self.ntaps = int(steps.size)
        self.tap_oz = int(oz)

    def coarse_shape(self, gi):
        g = self.groups[gi]
        nb = g.nb // 2
        return ((g.nx + 1) // 2 + 2 * nb + 12, (g.nz + 1 - getattr(self, "tap_oz", 0)) // 2 + nb + 12)

    def taps(self):
        """[entry] list of [ntaps][nxpad'][nzpad'] arrays (2x-coarser restrictions)."""
        shapes = [self.coarse_shape(gi) for gi, _, _ in self.entries]
        flat = np.zeros(sum(self.ntaps * a * b for a, b in shapes), np.float32)
        check(load().swfwi_batch_taps(self.h, flat))
        out, o = [], 0
        for a, b in shapes:
            n = self.ntaps * a * b
            out.append(flat[o:o + n].reshape(self.ntaps, a, b))
            o += n
        return out

    def step_from(self, prev, cur, it0, nsteps):
        """Fields [entry][nxpad][nzpad] (flat or stacked, entry order); returns
        (prev, cur) after `nsteps` steps from step it0."""
        prev = _f(prev).ravel()
        cur = _f(cur).ravel()
        po = np.zeros_like(prev)
        co = np.zeros_like(cur)
        check(load().swfwi_batch_step_from(self.h, prev, cur, int(it0), int(nsteps),
                                           po.ctypes.data_as(C.c_void_p), co.ctypes.data_as(C.c_void_p)))
        return po, co

    def set_correction(self, net):
        check(load().swfwi_batch_set_correction(self.h, None if net is None else net.h))
        self._net = net  # keep it alive while the batch refers to it

    def info(self):
        out = np.zeros(6, np.float64)
        check(load().swfwi_batch_info(self.h, out))
        return {"entries": int(out[0]), "forward_points": out[1], "backward_points": out[2],
                "snapshot_bytes": int(out[3]), "forward_ms": out[4], "backward_ms": out[5]}


class Net:
    """The learned coarse-grid correction network (swfwi_net): 3x3 convolutions
    2 -> 16 -> 16 -> 1, ReLU between, no biases; inputs p and r p (r = 1 /
    vel_t); y = net(p) is added to p."""

    NPARAMS = 9 * 2 * 16 + 9 * 16 * 16 + 9 * 16

    def __init__(self, weights=None, seed=0, scale=None):
        h = _VP()
        check(load().swfwi_net_create(C.byref(h)))
        self.h = h
        if weights is None:
            weights = self.initial(seed, scale)
        self.set_weights(weights)

    @staticmethod
    def initial(seed=0, scale=None):
        """He-style normal weights for the two ReLU layers, a small last layer
        (the network starts close to the identity correction y = 0)."""
        rng = np.random.default_rng(seed)
        w1 = rng.standard_normal(9 * 2 * 16) * np.sqrt(2.0 / 18)
        w2 = rng.standard_normal(9 * 16 * 16) * np.sqrt(2.0 / (9 * 16))
        w3 = rng.standard_normal(9 * 16) * (scale if scale is not None else 1e-3)
        return np.concatenate([w1, w2, w3]).astype(np.float32)

    def close(self):
        if getattr(self, "h", None):
            load().swfwi_net_destroy(self.h)
            self.h = None

    __del__ = close

    def set_weights(self, w):
        w = _f(w).ravel()
        assert w.size == self.NPARAMS
        check(load().swfwi_net_set_weights(self.h, w))

    def weights(self):
        w = np.zeros(self.NPARAMS, np.float32)
        check(load().swfwi_net_get_weights(self.h, w))
        return w

    def apply(self, fields, rv, region):
        f, r = _f(fields), _f(rv)
        e, nx, nz = f.shape
        assert r.shape == f.shape
        y = np.zeros_like(f)
        check(load().swfwi_net_apply(self.h, f, r, nx, nz, e, np.ascontiguousarray(region, np.int32), y))
        return y

    def time_apply(self, fields, rv, region, reps=20):
        """Milliseconds per launch of the forward kernel over device copies
        of the fields (a measurement probe)."""
        f, r = _f(fields), _f(rv)
        e, nx, nz = f.shape
        ms = C.c_double()
        check(load().swfwi_net_time_apply(self.h, f, r, nx, nz, e, np.ascontiguousarray(region, np.int32), reps,
                                          C.byref(ms)))
        return ms.value

    def loss_grad(self, inp, rv, target, region):
        a, r, t = _f(inp), _f(rv), _f(target)
        e, nx, nz = a.shape
        g = np.zeros(self.NPARAMS, np.float32)
        loss = C.c_double(0)
        check(load().swfwi_net_loss_grad(self.h, a, r, t, nx, nz, e, np.ascontiguousarray(region, np.int32),
                                         g.ctypes.data_as(C.c_void_p), C.byref(loss)))
        return loss.value, g

    def set_data(self, inp, rv, target, region):
        a, r, t = _f(inp), _f(rv), _f(target)
        n, nx, nz = a.shape
        check(load().swfwi_net_set_data(self.h, a, r, t, nx, nz, n, np.ascontiguousarray(region, np.int32)))

    def train(self, batch, steps, lr, seed=0):
        losses = np.zeros(steps, np.float64)
        check(load().swfwi_net_train(self.h, int(batch), int(steps), float(lr), int(seed),
                                     losses.ctypes.data_as(C.c_void_p)))
        return losses
