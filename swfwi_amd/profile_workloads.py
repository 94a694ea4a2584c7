"""Short, fixed workloads for rocprofv3 runs (kernel trace and HBM counters).

    python -m swfwi_amd.profile_workloads copy      # 256 Mi-float device copy (counter calibration)
    python -m swfwi_amd.profile_workloads shots     # 60-shot batched forward, 301x921, nt=400
    python -m swfwi_amd.profile_workloads ess       # one ESS epoch, 301x921, nt=1000
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
    case = workloads.config1(nt=nt)
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


def run_ess(nt, epochs):
    case = workloads.config1(nt=nt)
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
    for it in range(epochs):
        inv.ess_epoch(it)
        st = inv.stats()
        print(f"ess epoch {it}: obj {st['obj1']:.4e} grad {st['grad_ms']:.1f} ms "
              f"line search {st['linesearch_ms']:.1f} ms trials {st['ntrial']:.0f}", flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["copy", "shots", "ess"])
    ap.add_argument("--n", type=int, default=1 << 28)
    ap.add_argument("--nt", type=int, default=None)
    ap.add_argument("--epochs", type=int, default=2)
    a = ap.parse_args(argv)
    if capi.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        return 1
    if a.mode == "copy":
        run_copy(a.n)
    elif a.mode == "shots":
        run_shots(a.nt or 400)
    else:
        run_ess(a.nt or 1000, a.epochs)
    return 0


if __name__ == "__main__":
    np.seterr(all="ignore")
    sys.exit(main())
