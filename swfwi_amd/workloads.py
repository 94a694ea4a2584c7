"""Seeded synthetic workloads: layered velocity models and acquisition
geometries of the benchmark configurations, plus small cases for tests.

Host-side setup only (numpy/scipy); nothing here is on the GPU path."""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Case:
    name: str
    nz: int
    nx: int
    dx: float
    nb: int
    dt: float
    nt: int
    fm: float
    sx: np.ndarray
    sz: np.ndarray
    gx: np.ndarray
    gz: np.ndarray
    vmin: float
    vmax: float
    amp: float = 1.0
    seed: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def ns(self):
        return int(self.sx.size)

    @property
    def ng(self):
        return int(self.gx.size)

    @property
    def nxpad(self):
        return self.nx + 2 * self.nb + 12

    @property
    def nzpad(self):
        return self.nz + self.nb + 12

    @property
    def update_points(self):
        """Grid points one stencil step updates: (nxpad-12)(nzpad-12)."""
        return (self.nxpad - 12) * (self.nzpad - 12)


def layered_model(nz, nx, vmin, vmax, seed, nlayers=8, jump=300.0, rand_frac=0.05, smooth=4.0):
    """Depth gradient vmin->vmax + `nlayers` random dipping interfaces with
    +-`jump` m/s steps + a +-rand_frac Gaussian-smoothed random field.
    Returned as [nx][nz] float32 (z contiguous, the reference's layout)."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    z = np.arange(nz, dtype=np.float64)[:, None]
    x = np.arange(nx, dtype=np.float64)[None, :]
    v = vmin + (vmax - vmin) * z / max(nz - 1, 1) + 0.0 * x
    for _ in range(nlayers):
        z0 = rng.uniform(0.1, 0.95) * nz
        dip = rng.uniform(-0.15, 0.15)
        dv = rng.uniform(-jump, jump)
        v = v + dv * (z >= z0 + dip * (x - nx / 2.0))
    f = gaussian_filter(rng.standard_normal((nz, nx)), smooth)
    f = f / max(np.abs(f).max(), 1e-12) * rand_frac
    v = v * (1.0 + f)
    v = np.clip(v, vmin, vmax)
    return np.ascontiguousarray(v.T, np.float32)


def smoothed(v, sigma):
    from scipy.ndimage import gaussian_filter

    return np.ascontiguousarray(gaussian_filter(v.astype(np.float64), sigma), np.float32)


def fwi_201x301(nt=2000):
    """Plain 2-D acoustic FWI: 201x301 at 10 m, nb=20, dt=1 ms, 11 Ricker
    15 Hz sources at depth 1 cell, a receiver at every surface point."""
    nx, nz = 301, 201
    sx = np.arange(0, 301, 30, dtype=np.int32)
    return Case("fwi_201x301", nz, nx, 10.0, 20, 0.001, nt, 15.0, sx, np.ones_like(sx),
                np.arange(nx, dtype=np.int32), np.zeros(nx, np.int32), 1500.0, 4500.0, seed=0)


def ess_301x921(nt=4000):
    """ESS-FWI: 301x921 at 12.5 m, nb=30, dt=1 ms, 60 Ricker 10 Hz shots,
    921 surface receivers."""
    nx, nz = 921, 301
    sx = np.arange(0, 900, 15, dtype=np.int32)
    return Case("ess_301x921", nz, nx, 12.5, 30, 0.001, nt, 10.0, sx, np.ones_like(sx),
                np.arange(nx, dtype=np.int32), np.zeros(nx, np.int32), 1500.0, 5500.0, seed=1)


def small_case(nx=64, nz=40, nb=8, nt=420, ns=3, fm=25.0, dx=10.0, seed=5):
    """A test-sized acquisition: ns sources at depth 1, receivers on every
    surface cell; nt > 300 so the imaging condition is exercised."""
    sx = np.linspace(2, nx - 3, ns).astype(np.int32)
    return Case(f"small{nx}x{nz}", nz, nx, dx, nb, 0.001, nt, fm, sx, np.ones_like(sx),
                np.arange(nx, dtype=np.int32), np.zeros(nx, np.int32), 1500.0, 3500.0, seed=seed)


def models(case, sigma=None):
    """(true, initial) physical models [nx][nz] in m/s for a case."""
    v = layered_model(case.nz, case.nx, case.vmin, case.vmax, case.seed)
    s = sigma if sigma is not None else (20.0 if case.nx >= 200 else 4.0)
    return v, smoothed(v, s)


def ensemble_members(case, v0, N, seed=2, corr=20.0, std=0.03):
    """Ensemble-Kalman FWI members: v0 times (1 + a seeded Gaussian random
    field of correlation length `corr` cells and standard deviation `std`),
    expanded to the padded grid in transformed units, [N][nxpad][nzpad]."""
    from scipy.ndimage import gaussian_filter

    from . import capi

    rng = np.random.default_rng(seed)
    mem = []
    for _ in range(N):
        f = gaussian_filter(rng.standard_normal(v0.shape), corr)
        f = f / f.std() * std
        mem.append(capi.expand_domain((v0 * (1 + f)).astype(np.float32), case.nb, case.dx, case.dt))
    return np.stack(mem)


def add_noise(dobs, frac=0.01, seed=2):
    """Observed data plus seeded Gaussian noise at `frac` of its RMS amplitude."""
    rng = np.random.default_rng(seed)
    rms = float(np.sqrt(np.mean(dobs.astype(np.float64) ** 2)))
    return (dobs + rng.standard_normal(dobs.shape) * frac * rms).astype(np.float32)
