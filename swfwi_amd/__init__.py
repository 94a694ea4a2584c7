"""swfwi_amd -- MI355X-native acoustic FD propagation and adjoint-gradient
path of swfwi (2-D acoustic FWI), behind a C ABI (include/swfwi_hip.h).

Python here is a thin ctypes layer for tests and benchmarks; the propagator,
the gradient and the inversion drivers run in libswfwi_hip.so."""
from . import capi  # noqa: F401
from .capi import PREC_EXACT, PREC_FP32, Inversion, Solver, SwfwiError  # noqa: F401

__all__ = ["capi", "Solver", "Inversion", "SwfwiError", "PREC_EXACT", "PREC_FP32"]
