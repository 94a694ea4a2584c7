"""Python handle of the GPU ensemble Kalman analysis (swfwi_enkf)."""
import ctypes as C

import numpy as np

from .capi import _I, _VP, _f, check, load


class Enkf:
    """EnkfAnalyze over a capi.Solver: N velocity members on the device."""

    def __init__(self, solver, wlt, dobs, sigfac, nmembers):
        self.solver = solver
        self.N = int(nmembers)
        h = _VP()
        check(load().swfwi_enkf_create(C.byref(h), solver.h, _f(wlt), _f(dobs), sigfac, self.N))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().swfwi_enkf_destroy(self.h)
            self.h = None

    __del__ = close

    def set_members(self, vels):
        vels = _f(vels)
        assert vels.shape == (self.N,) + self.solver.shape
        check(load().swfwi_enkf_set_members(self.h, vels))

    def members(self):
        v = np.zeros((self.N,) + self.solver.shape, np.float32)
        check(load().swfwi_enkf_get_members(self.h, v))
        return v

    def analyze(self, perturbation=None, seed=0):
        if perturbation is None:
            check(load().swfwi_enkf_analyze(self.h, None, seed))
        else:
            e = np.ascontiguousarray(perturbation, np.float64)
            assert e.shape == (self.N, self.solver.nt, self.solver.ng)
            check(load().swfwi_enkf_analyze(self.h, e.ctypes.data_as(C.c_void_p), seed))

    def mean(self):
        m = np.zeros(self.solver.shape, np.float32)
        check(load().swfwi_enkf_mean(self.h, m))
        return m

    def info(self):
        sig = C.c_double(0)
        clip = _I(0)
        sv = np.zeros(self.N, np.float64)
        gain = np.zeros((self.N, self.N), np.float64)
        check(load().swfwi_enkf_info(self.h, C.byref(sig), C.byref(clip), sv.ctypes.data_as(C.c_void_p),
                                     gain.ctypes.data_as(C.c_void_p)))
        # gain is column major t4[j + N*i]: as a numpy (i, j) array -> transpose
        return {"sigma": sig.value, "clip": clip.value, "svals": sv, "gain": gain.T.copy()}


class EnsembleFwi:
    """The enfwi-damp driver over an Enkf: one EssFwiFramework per member,
    every member's epoch batched (swfwi_enfwi)."""

    KEYS = ["obj1", "steplen", "alpha2", "alpha3", "obj2", "obj3", "updateobj", "initobj", "ntrial", "parabolic"]

    def __init__(self, enkf, vmin, vmax, nita, maxdv):
        self.enkf = enkf
        self.solver = enkf.solver
        h = _VP()
        check(load().swfwi_enfwi_create(C.byref(h), self.solver.h, enkf.h, vmin, vmax, int(nita), maxdv))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().swfwi_enfwi_destroy(self.h)
            self.h = None

    __del__ = close

    def epoch(self, it):
        check(load().swfwi_enfwi_epoch(self.h, it))

    def mean_objective(self, want_mean=False):
        obj = C.c_float(0)
        if want_mean:
            m = np.zeros(self.solver.shape, np.float32)
            check(load().swfwi_enfwi_mean_objective(self.h, m.ctypes.data_as(C.c_void_p), C.byref(obj)))
            return obj.value, m
        check(load().swfwi_enfwi_mean_objective(self.h, None, C.byref(obj)))
        return obj.value

    def stats(self):
        out = np.zeros((self.enkf.N, 10), np.float32)
        ms = np.zeros(2, np.float64)
        cnt = np.zeros(2, np.int32)
        check(load().swfwi_enfwi_stats(self.h, out, ms, cnt))
        per = [dict(zip(self.KEYS, row.tolist())) for row in out]
        return {"members": per, "grad_ms": float(ms[0]), "linesearch_ms": float(ms[1]), "rounds": int(cnt[0]),
                "trials": int(cnt[1])}

    def gradients(self):
        g = np.zeros((self.enkf.N,) + self.solver.shape, np.float32)
        check(load().swfwi_enfwi_gradients(self.h, g))
        return g
