"""ctypes binding of libswfwi_hip.so (include/swfwi_hip.h).

The library is built in-tree by swfwi_amd.build; this module refuses to work
without it (there is no CPU fallback of the product path)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libswfwi_hip.so")

PREC_EXACT = 0
PREC_FP32 = 1

SNAP_F32, SNAP_BF16, SNAP_F16, SNAP_NONE = 0, 1, 2, 3
ORDER_10, ORDER_4 = 10, 4

# status codes (include/swfwi_hip.h)
OK, ERR_ARG, ERR_STATE, ERR_HIP, ERR_NUMERIC, ERR_RANGE = 0, 1, 2, 3, 4, 5

_f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64 = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_VP = C.c_void_p
_I, _F = C.c_int, C.c_float

# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "swfwi_last_error": (C.c_char_p, []),
    "swfwi_version": (C.c_char_p, []),
    "swfwi_device_count": (_I, [C.POINTER(_I)]),
    "swfwi_copy_bandwidth": (_I, [C.c_longlong, _I, C.POINTER(C.c_double)]),
    "swfwi_launch_latency": (_I, [_I, _I, _I, C.POINTER(C.c_double)]),
    "fd4t10s_damp_zjh_2d_vtrans": (None, [_f32, _f32, _f32, _f32, _I, _I, _I]),
    "fd4t10s_zjh_2d_vtrans": (None, [_f32, _f32, _f32, _f32, _I, _I]),
    "swfwi_step_damped": (_I, [_f32, _f32, _f32, _f32, _I, _I, _I, _I]),
    "swfwi_step_undamped": (_I, [_f32, _f32, _f32, _f32, _I, _I, _I]),
    "swfwi_expand_domain": (_I, [_f32, _I, _I, _I, _F, _F, _f32]),
    "swfwi_ricker": (_I, [_f32, _I, _F, _F, _F]),
    "swfwi_random_codes": (_I, [_I, _I, _I, _i32]),
    "swfwi_solver_create": (_I, [C.POINTER(_VP), _I, _I, _I, _F, _F, _F, _I, _I, _i32, _i32, _I, _i32, _i32, _I]),
    "swfwi_solver_destroy": (_I, [_VP]),
    "swfwi_solver_dims": (_I, [_VP, _i32]),
    "swfwi_solver_set_precision": (_I, [_VP, _I]),
    "swfwi_solver_bind_velocity": (_I, [_VP, _f32]),
    "swfwi_solver_get_velocity": (_I, [_VP, _f32]),
    "swfwi_forward_modeling": (_I, [_VP, _f32, _I, _f32]),
    "swfwi_forward_modeling_noalloc": (_I, [_VP, _f32, _I, _f32]),
    "swfwi_write_bndry": (_I, [_VP, _f32, _f32]),
    "swfwi_read_bndry": (_I, [_VP, _f32, _f32]),
    "swfwi_calgradient": (_I, [_VP, _f32, _I, _f32, _f32]),
    "swfwi_forward_shots": (_I, [_VP, _f32, _f32]),
    "swfwi_solver_counters": (_I, [_VP, _f64, np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")]),
    "swfwi_solver_set_graphs": (_I, [_VP, _I]),
    "swfwi_solver_set_tuning": (_I, [_VP, _I]),
    "swfwi_solver_set_tile_x": (_I, [_VP, _I]),
    "swfwi_solver_get_tile_x": (_I, [_VP, _I, _I, C.POINTER(_I)]),
    "swfwi_solver_set_timing": (_I, [_VP, _I]),
    "swfwi_solver_kernel_times": (_I, [_VP, _f64, np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")]),
    "swfwi_solver_time_steps": (_I, [_VP, _I, _I, _I, _I, C.POINTER(_F)]),
    "swfwi_solver_set_bwd_split": (_I, [_VP, _I]),
    "swfwi_enkf_create": (_I, [C.POINTER(_VP), _VP, _f32, _f32, _F, _I]),
    "swfwi_enkf_destroy": (_I, [_VP]),
    "swfwi_enkf_set_members": (_I, [_VP, _f32]),
    "swfwi_enkf_get_members": (_I, [_VP, _f32]),
    "swfwi_enkf_analyze": (_I, [_VP, C.c_void_p, C.c_ulonglong]),
    "swfwi_enkf_mean": (_I, [_VP, _f32]),
    "swfwi_enkf_info": (_I, [_VP, C.POINTER(C.c_double), C.POINTER(_I), C.c_void_p, C.c_void_p]),
    "swfwi_enfwi_create": (_I, [C.POINTER(_VP), _VP, _VP, _F, _F, _I, _F]),
    "swfwi_enfwi_destroy": (_I, [_VP]),
    "swfwi_enfwi_epoch": (_I, [_VP, _I]),
    "swfwi_enfwi_mean_objective": (_I, [_VP, C.c_void_p, C.POINTER(_F)]),
    "swfwi_enfwi_stats": (_I, [_VP, _f32, _f64, _i32]),
    "swfwi_enfwi_gradients": (_I, [_VP, _f32]),
    "swfwi_batch_create": (_I, [C.POINTER(_VP), _I, _i32, _f32, _i32, _i32, _I, _I]),
    "swfwi_batch_set_taps": (_I, [_VP, _I, _i32, _I]),
    "swfwi_batch_taps": (_I, [_VP, _f32]),
    "swfwi_batch_step_from": (_I, [_VP, _f32, _f32, _I, _I, C.c_void_p, C.c_void_p]),
    "swfwi_batch_set_correction": (_I, [_VP, _VP]),
    "swfwi_net_create": (_I, [C.POINTER(_VP)]),
    "swfwi_net_destroy": (_I, [_VP]),
    "swfwi_net_set_weights": (_I, [_VP, _f32]),
    "swfwi_net_get_weights": (_I, [_VP, _f32]),
    "swfwi_net_apply": (_I, [_VP, _f32, _f32, _I, _I, _I, _i32, _f32]),
    "swfwi_net_time_apply": (_I, [_VP, _f32, _f32, _I, _I, _I, _i32, _I, C.POINTER(C.c_double)]),
    "swfwi_net_loss_grad": (_I, [_VP, _f32, _f32, _f32, _I, _I, _I, _i32, C.c_void_p, C.POINTER(C.c_double)]),
    "swfwi_net_set_data": (_I, [_VP, _f32, _f32, _f32, _I, _I, _I, _i32]),
    "swfwi_net_train": (_I, [_VP, _I, _I, _F, C.c_ulonglong, C.c_void_p]),
    "swfwi_batch_destroy": (_I, [_VP]),
    "swfwi_batch_set_snapshot_precision": (_I, [_VP, _I]),
    "swfwi_batch_set_graphs": (_I, [_VP, _I]),
    "swfwi_batch_set_tuning": (_I, [_VP, _I]),
    "swfwi_batch_tile_x": (_I, [_VP, C.POINTER(_I)]),
    "swfwi_batch_set_tile_x": (_I, [_VP, _I]),
    "swfwi_batch_set_block_order": (_I, [_VP, _I]),
    "swfwi_batch_block_order": (_I, [_VP, C.POINTER(_I)]),
    "swfwi_batch_bind_model": (_I, [_VP, _I, _I, _f32]),
    "swfwi_batch_forward": (_I, [_VP, C.c_void_p]),
    "swfwi_batch_adjoint_source": (_I, [_VP, _f32, C.c_void_p]),
    "swfwi_batch_backward": (_I, [_VP, C.c_void_p, C.c_void_p, C.c_void_p]),
    "swfwi_batch_info": (_I, [_VP, _f64]),
    "swfwi_inversion_create": (_I, [C.POINTER(_VP), _VP, _f32, _f32, _F, _F, _I, _F]),
    "swfwi_inversion_destroy": (_I, [_VP]),
    "swfwi_ess_epoch": (_I, [_VP, _I]),
    "swfwi_fwi_epoch": (_I, [_VP, _I]),
    "swfwi_inversion_stats": (_I, [_VP, _f32, _f64]),
    "swfwi_inversion_gradient": (_I, [_VP, _f32]),
    "swfwi_inversion_direction": (_I, [_VP, _f32]),
}

_LIB = None


class SwfwiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"swfwi error {code}: {msg}")
        self.code = code


def load():
    """Load the library (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m swfwi_amd.build` "
                               "(the HIP path has no fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc != 0:
        raise SwfwiError(rc, load().swfwi_last_error().decode())


def device_count():
    n = _I(0)
    check(load().swfwi_device_count(C.byref(n)))
    return n.value


def copy_bandwidth(n=1 << 28, iters=20):
    """GB/s (read + write) of a plain device copy of n floats."""
    g = C.c_double(0)
    check(load().swfwi_copy_bandwidth(n, iters, C.byref(g)))
    return g.value


def launch_latency(blocks=738, reps=200, graph=False):
    """Microseconds per back-to-back launch of a near-empty kernel."""
    us = C.c_double(0)
    check(load().swfwi_launch_latency(int(blocks), int(reps), int(bool(graph)), C.byref(us)))
    return us.value


def _f(a):
    return np.ascontiguousarray(a, np.float32)


def _i(a):
    return np.ascontiguousarray(a, np.int32)


def expand_domain(v, nb, dx, dt):
    v = _f(v)
    nx, nz = v.shape
    out = np.zeros((nx + 2 * nb + 12, nz + nb + 12), np.float32)
    check(load().swfwi_expand_domain(v, nx, nz, nb, dx, dt, out))
    return out


def ricker(nt, fm, dt, amp):
    w = np.zeros(nt, np.float32)
    check(load().swfwi_ricker(w, nt, fm, dt, amp))
    return w


def random_codes(seed, n, skip=0):
    c = np.zeros(n, np.int32)
    check(load().swfwi_random_codes(seed, skip, n, c))
    return c


def step_damped(prev, cur, vel, u2, nb, precision=PREC_EXACT):
    nx, nz = prev.shape
    check(load().swfwi_step_damped(prev, _f(cur), _f(vel), u2, nx, nz, nb, precision))


def step_undamped(prev, cur, vel, u2, precision=PREC_EXACT):
    nx, nz = prev.shape
    check(load().swfwi_step_undamped(prev, _f(cur), _f(vel), u2, nx, nz, precision))


class Solver:
    """Damp4t10d on the GPU: geometry + bound model (swfwi_solver)."""

    def __init__(self, nx, nz, nb, dt, dx, fm, nt, sx, sz, gx, gz, precision=PREC_EXACT):
        self.sx, self.sz, self.gx, self.gz = _i(sx), _i(sz), _i(gx), _i(gz)
        self.nt, self.ns, self.ng = nt, self.sx.size, self.gx.size
        self.dt, self.dx, self.fm, self.nb = dt, dx, fm, nb
        h = _VP()
        check(load().swfwi_solver_create(C.byref(h), nx, nz, nb, dt, dx, fm, nt, self.ns, self.sx, self.sz,
                                         self.ng, self.gx, self.gz, precision))
        self.h = h
        d = np.zeros(6, np.int32)
        check(load().swfwi_solver_dims(self.h, d))
        self.nxpad, self.nzpad, self.bx0, self.bz0, self.bndr_size, self.ntiles = (int(x) for x in d)

    def close(self):
        if getattr(self, "h", None):
            load().swfwi_solver_destroy(self.h)
            self.h = None

    __del__ = close

    @property
    def shape(self):
        return (self.nxpad, self.nzpad)

    def set_precision(self, p):
        check(load().swfwi_solver_set_precision(self.h, p))

    def bind_velocity(self, vel_t):
        vel_t = _f(vel_t)
        assert vel_t.shape == self.shape
        check(load().swfwi_solver_bind_velocity(self.h, vel_t))

    def velocity(self):
        v = np.zeros(self.shape, np.float32)
        check(load().swfwi_solver_get_velocity(self.h, v))
        return v

    def write_bndry(self, field):
        """Damp4t10d::writeBndry of one step: the field's boundary band."""
        b = np.zeros(self.bndr_size, np.float32)
        check(load().swfwi_write_bndry(self.h, _f(field), b))
        return b

    def read_bndry(self, bndr, field):
        """Damp4t10d::readBndry: `field` with its band replaced by `bndr`."""
        out = _f(field).copy()
        check(load().swfwi_read_bndry(self.h, _f(bndr), out))
        return out

    def forward_modeling(self, encsrc, shot=-1):
        dcal = np.zeros((self.nt, self.ng), np.float32)
        check(load().swfwi_forward_modeling(self.h, _f(encsrc), shot, dcal))
        return dcal

    def forward_modeling_noalloc(self, encsrc, shot=-1):
        """forward_modeling over the library's fixed per-GPU workspace (no
        device allocation); raises SwfwiError (code 5) past its size."""
        dcal = np.zeros((self.nt, self.ng), np.float32)
        check(load().swfwi_forward_modeling_noalloc(self.h, _f(encsrc), shot, dcal))
        return dcal

    def calgradient(self, encsrc, vsrc, shot=-1):
        g = np.zeros(self.shape, np.float32)
        check(load().swfwi_calgradient(self.h, _f(encsrc), shot, _f(vsrc), g))
        return g

    def forward_shots(self, wlt):
        dobs = np.zeros((self.ns, self.nt, self.ng), np.float32)
        check(load().swfwi_forward_shots(self.h, _f(wlt), dobs))
        return dobs

    def counters(self):
        """([forward, backward] point updates, [forward, backward] launches)."""
        pu = np.zeros(2, np.float64)
        la = np.zeros(2, np.int64)
        check(load().swfwi_solver_counters(self.h, pu, la))
        return pu.tolist(), la.tolist()

    def set_graphs(self, on):
        """False/0 off, True/1 batches of up to 8 entries, 2 every batch size."""
        check(load().swfwi_solver_set_graphs(self.h, int(on)))

    def set_tuning(self, on):
        check(load().swfwi_solver_set_tuning(self.h, int(bool(on))))

    def set_tile_x(self, tx):
        check(load().swfwi_solver_set_tile_x(self.h, int(tx)))

    def tile_x(self, kind, entries):
        t = _I(0)
        check(load().swfwi_solver_get_tile_x(self.h, kind, entries, C.byref(t)))
        return t.value

    def set_bwd_split(self, mode):
        """-1: by batch size (default), 0: fused backward blocks, 1: split."""
        check(load().swfwi_solver_set_bwd_split(self.h, int(mode)))

    def time_steps(self, kind, entries, tile_x, reps=100):
        """Microseconds per launch of back-to-back step kernels on scratch fields."""
        us = _F(0)
        check(load().swfwi_solver_time_steps(self.h, kind, entries, tile_x, reps, C.byref(us)))
        return us.value

    def set_timing(self, on):
        check(load().swfwi_solver_set_timing(self.h, int(bool(on))))

    def kernel_times(self):
        """{'fwd': (ms_total, launches), 'bwd': (...)} of the timed time loops."""
        ms = np.zeros(2, np.float64)
        n = np.zeros(2, np.int64)
        check(load().swfwi_solver_kernel_times(self.h, ms, n))
        return {"fwd": (float(ms[0]), int(n[0])), "bwd": (float(ms[1]), int(n[1]))}


class Inversion:
    """FwiFramework / EssFwiFramework over a Solver (swfwi_inversion)."""

    KEYS = ["obj1", "steplen", "alpha2", "alpha3", "obj2", "obj3", "updateobj", "initobj", "ntrial", "parabolic"]

    def __init__(self, solver, wlt, dobs, vmin, vmax, nita, maxdv):
        self.solver = solver
        h = _VP()
        check(load().swfwi_inversion_create(C.byref(h), solver.h, _f(wlt), _f(dobs), vmin, vmax, nita, maxdv))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().swfwi_inversion_destroy(self.h)
            self.h = None

    __del__ = close

    def ess_epoch(self, it):
        check(load().swfwi_ess_epoch(self.h, it))

    def fwi_epoch(self, it):
        check(load().swfwi_fwi_epoch(self.h, it))

    def stats(self):
        s = np.zeros(10, np.float32)
        ms = np.zeros(2, np.float64)
        check(load().swfwi_inversion_stats(self.h, s, ms))
        d = dict(zip(self.KEYS, s.tolist()))
        d["grad_ms"], d["linesearch_ms"] = ms.tolist()
        return d

    def gradient(self):
        g = np.zeros(self.solver.shape, np.float32)
        check(load().swfwi_inversion_gradient(self.h, g))
        return g

    def direction(self):
        g = np.zeros(self.solver.shape, np.float32)
        check(load().swfwi_inversion_direction(self.h, g))
        return g
