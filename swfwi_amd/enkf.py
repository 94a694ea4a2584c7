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
