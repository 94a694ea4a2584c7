"""Short, fixed workloads for rocprofv3 runs (kernel trace and HBM counters).

    python -m swfwi_amd.profile_workloads copy      # 256 Mi-float device copy (counter calibration)
    python -m swfwi_amd.profile_workloads shots     # 60-shot batched forward, 301x921, nt=400
    python -m swfwi_amd.profile_workloads ess       # ESS epochs, 301x921, nt=1000 (--no-graphs for counters)
    python -m swfwi_amd.profile_workloads fwi       # plain FWI epochs, 201x301, 11 shots, nt=2000
    python -m swfwi_amd.profile_workloads enkf      # EnKF analyses, 301x921, N members
    python -m swfwi_amd.profile_workloads enfwi     # ensemble-Kalman FWI epochs, 301x921, N members
    python -m swfwi_amd.profile_workloads sweep     # step-kernel time per tile width and batch size
    python -m swfwi_amd.profile_workloads net       # network apply kernel over 16 coarse fields of the 501x1501 case
"""
import argparse
import sys
import time

import numpy as np

from . import capi, workloads


def run_copy(n):
    gbs = capi.copy_bandwidth(n, 5)
    print(f"copy n={n} floats ({n * 4 / 2**20:.0f} MiB each way): {gbs:.1f} GB/s", flush=True)


def run_shots(nt):
    case = workloads.ess_301x921(nt=nt)
    vtrue, _ = workloads.models(case)
    vt = capi.expand_domain(vtrue, case.nb, case.dx, case.dt)
    wlt = capi.ricker(case.nt, case.fm, case.dt, case.amp)
    s = capi.Solver(case.nx, case.nz, case.nb, case.dt, case.dx, case.fm, case.nt, case.sx, case.sz, case.gx,
                    case.gz)
    s.bind_velocity(vt)
    t0 = time.perf_counter()
    s.forward_shots(wlt)
    dt = time.perf_counter() - t0
    print(f"shots: {case.ns} x {nt} steps in {dt * 1e3:.1f} ms "
          f"({case.ns * nt * case.update_points / dt / 1e9:.1f} Gpts/s incl. transfer)", flush=True)


def run_sweep(entries, reps, precision):
    """Per-launch time of every step kernel width over batch sizes on the
    301x921 grid (back-to-back launches on scratch wavefields; the backward
    iteration includes the imaging condition into a scratch gradient; no
    sources, receivers or band)."""
    case = workloads.ess_301x921(nt=8)
    vtrue, _ = workloads.models(case)
    s = capi.Solver(case.nx, case.nz, case.nb, case.dt, case.dx, case.fm, case.nt, case.sx, case.sz, case.gx,
                    case.gz, precision)
    s.bind_velocity(capi.expand_domain(vtrue, case.nb, case.dx, case.dt))
    pts = case.update_points
    for blocks in (256, 738, 2952):
        print(f"empty kernel, {blocks} blocks: stream {capi.launch_latency(blocks, reps):.2f} us, "
              f"graph {capi.launch_latency(blocks, reps, graph=True):.2f} us per launch", flush=True)
    for E in entries:
        for kind, byts, split, label in ((0, 16.0, -1, "fwd"), (1, 40.0, 0, "bwd"), (1, 40.0, 1, "bwd/split")):
            s.set_bwd_split(split)
            row = []
            for tx in (8, 16, 32):
                s.time_steps(kind, E, tx, 4)  # warm
                us = s.time_steps(kind, E, tx, reps)
                row.append(f"tx{tx} {us:8.1f} us {E * pts * byts / (us * 1e3):7.0f} GB/s")
            print(f"{label:9s} E={E:4d}: " + " | ".join(row), flush=True)
        s.set_bwd_split(-1)


def run_ess(nt, epochs, graphs=True):
    case = workloads.ess_301x921(nt=nt)
    vtrue, v0 = workloads.models(case)
    vt = capi.expand_domain(vtrue, case.nb, case.dx, case.dt)
    vi = capi.expand_domain(v0, case.nb, case.dx, case.dt)
    wlt = capi.ricker(case.nt, case.fm, case.dt, case.amp)
    s = capi.Solver(case.nx, case.nz, case.nb, case.dt, case.dx, case.fm, case.nt, case.sx, case.sz, case.gx,
                    case.gz)
    s.bind_velocity(vt)
    dobs = s.forward_shots(wlt)
    s.bind_velocity(vi)
    inv = capi.Inversion(s, wlt, dobs, case.vmin, case.vmax, 5, 200.0)
    if not graphs:
        s.set_graphs(False)  # one dispatch per launch, for per-dispatch counters
    for it in range(epochs):
        inv.ess_epoch(it)
        st = inv.stats()
        print(f"ess epoch {it}: obj {st['obj1']:.4e} grad {st['grad_ms']:.1f} ms "
              f"line search {st['linesearch_ms']:.1f} ms trials {st['ntrial']:.0f}", flush=True)


def run_fwi(nt, epochs):
    """Plain FWI on the 201x301 grid (nb=20, 11 shots, 301 receivers): the
    observed data, then `epochs` FwiFramework epochs (every shot's gradient
    in one batch, the trial forward runs batched); gradient time per shot."""
    case = workloads.fwi_201x301(nt=nt)
    vtrue, v0 = workloads.models(case)
    s = capi.Solver(case.nx, case.nz, case.nb, case.dt, case.dx, case.fm, case.nt, case.sx, case.sz, case.gx,
                    case.gz)
    wlt = capi.ricker(case.nt, case.fm, case.dt, case.amp)
    s.bind_velocity(capi.expand_domain(vtrue, case.nb, case.dx, case.dt))
    dobs = s.forward_shots(wlt)
    s.bind_velocity(capi.expand_domain(v0, case.nb, case.dx, case.dt))
    inv = capi.Inversion(s, wlt, dobs, case.vmin, case.vmax, 5, 200.0)
    for it in range(epochs):
        pu0 = sum(s.counters()[0])
        t0 = time.perf_counter()
        inv.fwi_epoch(it)
        dt = time.perf_counter() - t0
        pu = sum(s.counters()[0]) - pu0
        st = inv.stats()
        print(f"fwi epoch {it}: {case.ns} shots, {dt * 1e3:.1f} ms ({pu / dt / 1e9:.1f} Gpts/s), gradient "
              f"{st['grad_ms']:.1f} ms ({st['grad_ms'] / case.ns:.2f} ms per shot), line search "
              f"{st['linesearch_ms']:.1f} ms, obj {st['obj1']:.4e}", flush=True)


def run_enkf(nt, members, graphs=True):
    """EnKF analysis on the 301x921 grid: `members` velocity members (initial
    model + 3% Gaussian random fields, correlation ~20 cells), one batched
    encoded forward run of all members and the FP64 analysis."""
    from scipy.ndimage import gaussian_filter

    from .enkf import Enkf

    case = workloads.ess_301x921(nt=nt)
    vtrue, v0 = workloads.models(case)
    rng = np.random.default_rng(2)
    mem = []
    for _ in range(members):
        f = gaussian_filter(rng.standard_normal(v0.shape), 20.0)
        f = f / f.std() * 0.03
        mem.append(capi.expand_domain((v0 * (1 + f)).astype(np.float32), case.nb, case.dx, case.dt))
    wlt = capi.ricker(case.nt, case.fm, case.dt, case.amp)
    s = capi.Solver(case.nx, case.nz, case.nb, case.dt, case.dx, case.fm, case.nt, case.sx, case.sz, case.gx,
                    case.gz)
    s.set_graphs(graphs)
    s.bind_velocity(capi.expand_domain(vtrue, case.nb, case.dx, case.dt))
    dobs = s.forward_shots(wlt)
    e = Enkf(s, wlt, dobs, 0.2, members)
    e.set_members(np.stack(mem))
    for k in range(2):
        pu0 = sum(s.counters()[0])
        t0 = time.perf_counter()
        e.analyze(None, seed=k)
        dt = time.perf_counter() - t0
        pu = sum(s.counters()[0]) - pu0
        info = e.info()
        print(f"enkf analysis {k}: N={members} nt={nt} {dt * 1e3:.1f} ms ({pu / dt / 1e9:.1f} Gpts/s) "
              f"clip={info['clip']} sigma={info['sigma']:.3e}", flush=True)


def run_enfwi(nt, members, epochs, graphs=True):
    """Ensemble-Kalman FWI on the 301x921 grid: `members` velocity members
    (initial model + 3% Gaussian random fields, correlation ~20 cells),
    observed data with 1% RMS Gaussian noise; the initial analysis, then
    `epochs` batched member epochs with an analysis after each."""
    from scipy.ndimage import gaussian_filter

    from .enkf import Enkf, EnsembleFwi

    case = workloads.ess_301x921(nt=nt)
    vtrue, v0 = workloads.models(case)
    rng = np.random.default_rng(2)
    mem = []
    for _ in range(members):
        f = gaussian_filter(rng.standard_normal(v0.shape), 20.0)
        f = f / f.std() * 0.03
        mem.append(capi.expand_domain((v0 * (1 + f)).astype(np.float32), case.nb, case.dx, case.dt))
    wlt = capi.ricker(case.nt, case.fm, case.dt, case.amp)
    s = capi.Solver(case.nx, case.nz, case.nb, case.dt, case.dx, case.fm, case.nt, case.sx, case.sz, case.gx,
                    case.gz)
    s.set_graphs(graphs)
    s.bind_velocity(capi.expand_domain(vtrue, case.nb, case.dx, case.dt))
    dobs = s.forward_shots(wlt)
    rms = float(np.sqrt(np.mean(dobs.astype(np.float64) ** 2)))
    dobs = (dobs + rng.standard_normal(dobs.shape) * 0.01 * rms).astype(np.float32)
    e = Enkf(s, wlt, dobs, 0.2, members)
    e.set_members(np.stack(mem))
    fwi = EnsembleFwi(e, case.vmin, case.vmax, 5, 200.0)
    t0 = time.perf_counter()
    e.analyze(None, seed=0)
    print(f"initial analysis {1e3 * (time.perf_counter() - t0):.0f} ms, mean-model objective "
          f"{fwi.mean_objective():.4e}", flush=True)
    for it in range(epochs):
        pu0 = sum(s.counters()[0])
        t0 = time.perf_counter()
        fwi.epoch(it)
        t1 = time.perf_counter()
        pu = sum(s.counters()[0]) - pu0
        st = fwi.stats()
        e.analyze(None, seed=it + 1)
        t2 = time.perf_counter()
        steps = [m["steplen"] for m in st["members"]]
        print(f"epoch {it}: N={members} {1e3 * (t1 - t0):.0f} ms ({pu / (t1 - t0) / 1e9:.1f} Gpts/s; gradients "
              f"{st['grad_ms']:.0f} ms, line search {st['linesearch_ms']:.0f} ms in {st['rounds']} rounds / "
              f"{st['trials']} trial runs), analysis {1e3 * (t2 - t1):.0f} ms, mean-model objective "
              f"{fwi.mean_objective():.4e}, steps {min(steps):.3g}..{max(steps):.3g}", flush=True)


def run_net(reps):
    """The learned correction's forward kernel alone: 16 smooth random fields
    of the coarse grid of the 501x1501 learned-FWI case (793 x 277 padded),
    1 / vel_t = 0.1, random weights; `reps` back-to-back launches, twice."""
    from scipy.ndimage import gaussian_filter

    from . import learned
    from .batch import Net

    case = learned.fwi_501x1501_case()
    _, coarse = learned.grids(case)
    shape = (coarse["nx"] + 2 * coarse["nb"] + 12, coarse["nz"] + coarse["nb"] + 12)
    rng = np.random.default_rng(1)
    fields = np.stack([gaussian_filter(rng.standard_normal(shape), 3.0) for _ in range(16)]).astype(np.float32)
    rv = np.full(fields.shape, 0.1, np.float32)
    reg = learned.region(coarse)
    net = Net(Net.initial(seed=3, scale=0.1))
    npts = (reg[1] - reg[0]) * (reg[3] - reg[2]) * 16
    for _ in range(2):
        ms = net.time_apply(fields, rv, reg, reps=reps)
        print(f"net apply {shape} x16: {1e3 * ms:.2f} us per launch, {5472 * npts / (ms * 1e-3) / 1e12:.1f} TFLOP/s",
              flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["copy", "shots", "ess", "fwi", "enkf", "enfwi", "sweep", "net"])
    ap.add_argument("--entries", default="1,2,3,4,8,16,60,100,200")
    ap.add_argument("--reps", type=int, default=100)
    ap.add_argument("--precision", choices=["exact", "fp32"], default="exact")
    ap.add_argument("--members", type=int, default=100)
    ap.add_argument("--n", type=int, default=1 << 28)
    ap.add_argument("--nt", type=int, default=None)
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--no-graphs", action="store_true", help="launch directly instead of replaying HIP graphs")
    a = ap.parse_args(argv)
    if capi.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        return 1
    if a.mode == "copy":
        run_copy(a.n)
    elif a.mode == "shots":
        run_shots(a.nt or 400)
    elif a.mode == "sweep":
        run_sweep([int(x) for x in a.entries.split(",")], a.reps,
                  capi.PREC_FP32 if a.precision == "fp32" else capi.PREC_EXACT)
    elif a.mode == "enfwi":
        run_enfwi(a.nt or 4000, a.members, a.epochs, graphs=not a.no_graphs)
    elif a.mode == "enkf":
        run_enkf(a.nt or 4000, a.members, graphs=not a.no_graphs)
    elif a.mode == "fwi":
        run_fwi(a.nt or 2000, a.epochs)
    elif a.mode == "net":
        run_net(a.reps)
    else:
        run_ess(a.nt or 1000, a.epochs, graphs=not a.no_graphs)
    return 0


if __name__ == "__main__":
    np.seterr(all="ignore")
    sys.exit(main())
