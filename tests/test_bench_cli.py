"""bench.py's host side on the CPU: the one-GPU contract, the committed HBM
traffic figure it reports, and the CPU-baseline leg (the oracle timed on a
small case, encoded and single-shot)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_measured_traffic_reads_the_committed_profile():
    import bench
    t = bench.measured_traffic("k_fwd_step", "ess_301x921")
    assert isinstance(t, int) and t > 0
    tb = bench.measured_traffic("k_fwd_step_batched", "ess_301x921")
    assert isinstance(tb, int) and tb > t
    assert bench.measured_traffic("k_fwd_step", "fwi_201x301") is None
    assert bench.measured_traffic("k_bwd_step", "ess_301x921") is None


@pytest.mark.parametrize("encoded", [True, False])
def test_cpu_baseline_on_a_small_case(oracle, encoded):
    import bench
    from swfwi_amd import workloads
    case = workloads.small_case(nx=40, nz=30, nb=8, nt=320, ns=3)
    out, threads = bench.cpu_baseline(case, encoded)
    assert threads >= 1
    for label in ("reference_threading", "all_parallel"):
        gpts, secs, nt, iters = out[label]
        assert gpts > 0 and secs > 0
        assert nt == case.nt
        # nt=320 at dt=1 ms: backward iterations it = 319..299 (the reference's
        # taper test passes from it=300; it=299 runs and breaks)
        assert iters == 21


def test_cpu_epoch_on_a_small_case(oracle):
    """The whole-epoch CPU baseline: one oracle ESS epoch (gradient + line
    search) on data the oracle itself made; it reports its trial runs."""
    import numpy as np

    import bench
    from swfwi_amd import workloads
    case = workloads.small_case(nx=40, nz=30, nb=8, nt=320, ns=3)
    vtrue, _ = workloads.models(case)
    g = oracle.Geom(case.nx, case.nz, case.nb, case.dt, case.dx, case.fm, case.nt, case.sx, case.sz, case.gx, case.gz)
    wlt = oracle.ricker(case.nt, case.fm, case.dt, case.amp)
    vt = oracle.expand_domain(vtrue, case.nb, case.dx, case.dt)
    dobs = np.stack([g.forward_modeling(vt, wlt, shot=s) for s in range(case.ns)])
    secs, ntrial = bench.cpu_epoch(case, True, dobs)
    assert secs > 0
    assert ntrial >= 1


def test_mixed_groups_are_the_mixed_size_workload():
    """4 models of 201x301 (nt=2000), 2 of 301x921 (nt=4000), 1 of 501x1501
    (nt=6000), 16 surface shots each inside the model, a receiver on every
    surface cell."""
    import bench
    groups = bench.mixed_groups(16)
    assert [(g.nz, g.nx, g.nt, g.nmodels) for g in groups] == [(201, 301, 2000, 4), (301, 921, 4000, 2),
                                                                (501, 1501, 6000, 1)]
    for g in groups:
        assert len(g.sx) == 16 and min(g.sx) == 0 and max(g.sx) == g.nx - 1
        assert g.ng == g.nx and g.gz == 0 and all(z == 1 for z in g.sz)
