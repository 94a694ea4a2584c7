"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the reference's
own stencil kernels compiled unchanged (oracle/_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the CPU-baseline leg of bench.py, as the results the GPU path is checked
against.  The product package never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
VP = C.c_void_p
I, F = C.c_int, C.c_float
OBJ_HOOK = C.CFUNCTYPE(C.c_float, C.c_float, C.c_void_p)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(path)
        sig = {
            "orc_set_sum_mode": (None, [I]),
            "orc_max_threads": (I, []),
            "orc_fd4t10s_damp": (None, [f32p, f32p, f32p, f32p, I, I, I, I]),
            "orc_fd4t10s": (None, [f32p, f32p, f32p, f32p, I, I]),
            "orc_expand_domain": (None, [f32p, I, I, I, F, F, f32p]),
            "orc_recover_vel": (None, [f32p, C.c_long, F, F]),
            "orc_vel_trans": (F, [F, F, F]),
            "orc_vel_recover": (F, [F, F, F]),
            "orc_geom_new": (VP, [I, I, I, F, F, F, I, I, i32p, i32p, I, i32p, i32p, I]),
            "orc_geom_free": (None, [VP]),
            "orc_geom_dims": (None, [VP, i32p]),
            "orc_geom_set_parallel_u2": (None, [VP, I]),
            "orc_bndr_size": (I, [VP]),
            "orc_write_bndr": (None, [VP, f32p, f32p]),
            "orc_read_bndr": (None, [VP, f32p, f32p]),
            "orc_add_sources": (None, [VP, f32p, f32p, I, I, I]),
            "orc_add_receivers": (None, [VP, f32p, f32p, I]),
            "orc_record_seis": (None, [VP, f32p, f32p]),
            "orc_forward_modeling": (None, [VP, f32p, f32p, I, f32p]),
            "orc_calgradient": (None, [VP, f32p, f32p, I, f32p, f32p]),
            "orc_xcorr_weight": (I, [F, I, C.POINTER(F)]),
            "orc_trans_vsrc": (None, [f32p, I, I]),
            "orc_mute_windows": (None, [VP, f32p, I, i32p, i32p]),
            "orc_remove_direct_arrival": (None, [VP, f32p, I, f32p]),
            "orc_cal_objective": (F, [f32p, C.c_long]),
            "orc_scale_gradient": (None, [VP, f32p]),
            "orc_mask_gradient": (None, [VP, f32p]),
            "orc_refill_boundary": (None, [VP, f32p]),
            "orc_update_grad": (None, [f32p, f32p, f32p, C.c_long, I]),
            "orc_vel_bounds": (None, [F, F, F, F, C.POINTER(F), C.POINTER(F)]),
            "orc_update_vel": (None, [f32p, f32p, f32p, C.c_long, F, F, F]),
            "orc_cal_max_alpha23": (None, [f32p, f32p, C.c_long, F, F, F, C.POINTER(F), C.POINTER(F)]),
            "orc_parabola_vertex": (None, [F, F, F, F, F, F, F, C.POINTER(F), C.POINTER(F)]),
            "orc_ricker": (None, [f32p, I, F, F, F]),
            "orc_encode_source": (None, [i32p, I, f32p, I, f32p]),
            "orc_encode_obs": (None, [i32p, I, f32p, I, I, f32p]),
            "orc_clip_position": (I, [f64p, I]),
            "orc_ess_new": (VP, [VP, f32p, f32p, f32p, F, F, I, F]),
            "orc_ess_delete": (None, [VP]),
            "orc_ess_epoch": (None, [VP, I]),
            "orc_fwi_epoch": (None, [VP, I]),
            "orc_ess_scalars": (None, [VP, f32p]),
            "orc_ess_copy_grad": (None, [VP, f32p]),
            "orc_ess_copy_dir": (None, [VP, f32p]),
            "orc_set_objective_hook": (None, [OBJ_HOOK, VP]),
            "o64_forward_modeling": (None, [I, I, I, I, F, f32p, f64p, I, I, I, i32p, i32p, f64p]),
            "o64_calgradient": (None, [I, I, I, I, F, f32p, f64p, I, I, I, i32p, i32p, f64p, f64p]),
            "orc_steplen_scripted": (None, [F, F, I, f32p, f32p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's fd4t10s kernels compiled unchanged, or None if the
    reference tree was not present when the oracle was built."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        L.fd4t10s_damp_zjh_2d_vtrans.restype = None
        L.fd4t10s_damp_zjh_2d_vtrans.argtypes = [f32p, f32p, f32p, f32p, I, I, I]
        L.fd4t10s_zjh_2d_vtrans.restype = None
        L.fd4t10s_zjh_2d_vtrans.argtypes = [f32p, f32p, f32p, f32p, I, I]
        _REF = L
    return _REF


def set_sum_mode(mode):
    lib().orc_set_sum_mode(int(mode))


# -- kernels ---------------------------------------------------------------

def fd4t10s_damp(prev, cur, vel, u2, nb, parallel_u2=1):
    nx, nz = prev.shape
    lib().orc_fd4t10s_damp(prev, cur, vel, u2, nx, nz, nb, parallel_u2)


def fd4t10s(prev, cur, vel, u2):
    nx, nz = prev.shape
    lib().orc_fd4t10s(prev, cur, vel, u2, nx, nz)


def expand_domain(v, nb, dx, dt):
    nx, nz = v.shape
    out = np.zeros((nx + 2 * nb + 12, nz + nb + 12), np.float32)
    lib().orc_expand_domain(np.ascontiguousarray(v, np.float32), nx, nz, nb, dx, dt, out)
    return out


def ricker(nt, fm, dt, amp):
    w = np.zeros(nt, np.float32)
    lib().orc_ricker(w, nt, fm, dt, amp)
    return w


def parabola_vertex(x1, y1, x2, y2, x3, y3, maxa3=0.0):
    xv, yv = F(), F()
    lib().orc_parabola_vertex(x1, y1, x2, y2, x3, y3, maxa3, C.byref(xv), C.byref(yv))
    return xv.value, yv.value


def vel_bounds(vmin, vmax, dx, dt):
    a, b = F(), F()
    lib().orc_vel_bounds(vmin, vmax, dx, dt, C.byref(a), C.byref(b))
    return a.value, b.value


def cal_max_alpha23(vel, grad, dt, dx, maxdv):
    a2, a3 = F(), F()
    lib().orc_cal_max_alpha23(vel.ravel(), grad.ravel(), vel.size, dt, dx, maxdv, C.byref(a2), C.byref(a3))
    return a2.value, a3.value


def steplen_scripted(objective, obj1, max_alpha3, nita, preserved=(0.0, 0)):
    """calsteplen's search (initAlpha23 + refineAlpha + parabola) with
    calobjval replaced by objective(alpha) -> float.  Returns (trial alphas in
    evaluation order, scalars as Ess.scalars, new preserved state)."""
    trials = []

    def hook(alpha, _ctx):
        trials.append(alpha)
        return float(np.float32(objective(alpha)))

    cb = OBJ_HOOK(hook)
    state = np.array([preserved[0], preserved[1]], np.float32)
    out = np.zeros(10, np.float32)
    lib().orc_set_objective_hook(cb, None)
    try:
        lib().orc_steplen_scripted(obj1, max_alpha3, nita, state, out)
    finally:
        lib().orc_set_objective_hook(OBJ_HOOK(), None)
    keys = ["obj1", "steplen", "alpha2", "alpha3", "obj2", "obj3", "updateobj", "initobj", "ntrial", "parabolic"]
    return trials, dict(zip(keys, out.tolist())), (float(state[0]), int(state[1]))


def clip_position(s):
    s = np.ascontiguousarray(s, np.float64)
    return lib().orc_clip_position(s, s.size)


class Geom:
    """Padded geometry of a Damp4t10d(allSrcPos, allGeoPos, dt, dx, fm, nb, nt)
    bound to an expanded model of physical size nx x nz."""

    def __init__(self, nx, nz, nb, dt, dx, fm, nt, sx, sz, gx, gz, parallel_u2=1):
        self.sx = np.ascontiguousarray(sx, np.int32)
        self.sz = np.ascontiguousarray(sz, np.int32)
        self.gx = np.ascontiguousarray(gx, np.int32)
        self.gz = np.ascontiguousarray(gz, np.int32)
        self.ns, self.ng = self.sx.size, self.gx.size
        self.nt, self.dt, self.dx, self.fm, self.nb = nt, dt, dx, fm, nb
        self.h = lib().orc_geom_new(nx, nz, nb, dt, dx, fm, nt, self.ns, self.sx, self.sz, self.ng, self.gx,
                                    self.gz, parallel_u2)
        d = np.zeros(5, np.int32)
        lib().orc_geom_dims(self.h, d)
        self.nxpad, self.nzpad, self.bx0, self.bz0, self.bndr_size = (int(v) for v in d)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_geom_free(self.h)
            self.h = None

    @property
    def shape(self):
        return (self.nxpad, self.nzpad)

    def zeros(self):
        return np.zeros(self.shape, np.float32)

    def forward_modeling(self, vel, encsrc, shot=-1):
        dcal = np.zeros((self.nt, self.ng), np.float32)
        lib().orc_forward_modeling(self.h, vel, np.ascontiguousarray(encsrc, np.float32), shot, dcal)
        return dcal

    def calgradient(self, vel, encsrc, vsrc, shot=-1, g0=None):
        g = self.zeros() if g0 is None else g0
        lib().orc_calgradient(self.h, vel, np.ascontiguousarray(encsrc, np.float32), shot,
                              np.ascontiguousarray(vsrc, np.float32), g)
        return g

    def forward_modeling64(self, vel, src, shot):
        """FP64 run of FwiForwardModeling for source `shot` (amplitudes src[nt])."""
        dcal = np.zeros((self.nt, self.ng), np.float64)
        lib().o64_forward_modeling(self.nxpad, self.nzpad, self.nb, self.nt, self.dt, np.ascontiguousarray(vel),
                                   np.ascontiguousarray(src, np.float64), int(self.sx[shot]), int(self.sz[shot]),
                                   self.ng, self.gx, self.gz, dcal)
        return dcal

    def calgradient64(self, vel, src, vsrc, shot):
        """FP64 run of calgradient for source `shot`: raw imaging condition."""
        g = np.zeros(self.shape, np.float64)
        lib().o64_calgradient(self.nxpad, self.nzpad, self.nb, self.nt, self.dt, np.ascontiguousarray(vel),
                              np.ascontiguousarray(src, np.float64), int(self.sx[shot]), int(self.sz[shot]), self.ng,
                              self.gx, self.gz, np.ascontiguousarray(vsrc, np.float64), g)
        return g

    def mute_windows(self, vel, src=0):
        st = np.zeros(self.ng, np.int32)
        en = np.zeros(self.ng, np.int32)
        lib().orc_mute_windows(self.h, vel, src, st, en)
        return st, en

    def remove_direct_arrival(self, vel, data, src=0):
        lib().orc_remove_direct_arrival(self.h, vel, src, data)

    def scale_gradient(self, g):
        lib().orc_scale_gradient(self.h, g)

    def mask_gradient(self, g):
        lib().orc_mask_gradient(self.h, g)

    def refill_boundary(self, a):
        lib().orc_refill_boundary(self.h, a)

    def write_bndr(self, p):
        b = np.zeros(self.bndr_size, np.float32)
        lib().orc_write_bndr(self.h, b, p)
        return b

    def read_bndr(self, b, p):
        lib().orc_read_bndr(self.h, b, p)

    def add_sources(self, p, src, lo=0, hi=None, sign=1):
        hi = self.ns if hi is None else hi
        lib().orc_add_sources(self.h, p, np.ascontiguousarray(src, np.float32), lo, hi, sign)

    def add_receivers(self, p, src, sign=1):
        lib().orc_add_receivers(self.h, p, np.ascontiguousarray(src, np.float32), sign)

    def record(self, p):
        s = np.zeros(self.ng, np.float32)
        lib().orc_record_seis(self.h, s, p)
        return s


def trans_vsrc(vsrc):
    nt, ng = vsrc.shape
    lib().orc_trans_vsrc(vsrc, nt, ng)


def cal_objective(d):
    d = np.ascontiguousarray(d, np.float32).ravel()
    return lib().orc_cal_objective(d, d.size)


def update_grad(pre, cur, d, it):
    lib().orc_update_grad(pre.ravel(), cur.ravel(), d.ravel(), pre.size, it)


def update_vel(v, grad, steplen, tmin, tmax):
    out = np.empty_like(v)
    lib().orc_update_vel(out.ravel(), v.ravel(), grad.ravel(), v.size, steplen, tmin, tmax)
    return out


def gen_codes(seed, n, skip=0):
    """ns +-1 codes of RandomCodes(seed) after `skip` earlier draws."""
    x = seed % 2147483647 or 1
    out = []
    for k in range(skip + n):
        x = (x * 48271) % 2147483647
        if k >= skip:
            out.append(((x - 1) // 1073741823) * 2 - 1)
    return np.array(out, np.int32)


def encode_source(code, wlt):
    code = np.ascontiguousarray(code, np.int32)
    enc = np.zeros((wlt.size, code.size), np.float32)
    lib().orc_encode_source(code, code.size, np.ascontiguousarray(wlt, np.float32), wlt.size, enc)
    return enc


def encode_obs(code, dobs):
    code = np.ascontiguousarray(code, np.int32)
    ns, nt, ng = dobs.shape
    enc = np.zeros((nt, ng), np.float32)
    lib().orc_encode_obs(code, ns, np.ascontiguousarray(dobs, np.float32), nt, ng, enc)
    return enc


class Ess:
    """EssFwiFramework / FwiFramework state over an oracle geometry."""

    def __init__(self, geom, vel, wlt, dobs, vmin, vmax, nita, maxdv):
        self.geom = geom
        self.vel = vel
        self.wlt = np.ascontiguousarray(wlt, np.float32)
        self.dobs = np.ascontiguousarray(dobs, np.float32)
        self.h = lib().orc_ess_new(geom.h, vel, self.wlt, self.dobs, vmin, vmax, nita, maxdv)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ess_delete(self.h)
            self.h = None

    def epoch(self, it):
        lib().orc_ess_epoch(self.h, it)

    def fwi_epoch(self, it):
        lib().orc_fwi_epoch(self.h, it)

    def scalars(self):
        s = np.zeros(10, np.float32)
        lib().orc_ess_scalars(self.h, s)
        keys = ["obj1", "steplen", "alpha2", "alpha3", "obj2", "obj3", "updateobj", "initobj", "ntrial", "parabolic"]
        return dict(zip(keys, s.tolist()))

    def grad(self):
        g = self.geom.zeros()
        lib().orc_ess_copy_grad(self.h, g)
        return g

    def direction(self):
        g = self.geom.zeros()
        lib().orc_ess_copy_dir(self.h, g)
        return g
