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
