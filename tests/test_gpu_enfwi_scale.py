"""Ensemble-Kalman FWI at its workload size: N=100 velocity members on the
301x921 grid (nb=30, 60 shots, 921 surface receivers), members = initial
model x (1 + 3% Gaussian random field of correlation length 20 cells),
observed data with 1% RMS Gaussian noise (main-enfwi-damp.cpp:559-605).

* reduced nt (400, so the oracle finishes in seconds): one batched epoch of
  all 100 members, three of them against the oracle's own per-member
  EssFwiFramework epochs (essfwiframework.cpp:191-239);
* full nt=4000: one batched epoch of all members with property checks,
  then the FP64 analysis (enkfanalyze.cpp:343-495) on an injected
  perturbation against numpy's SVD of the same band (singular values, clip,
  t4, updated members), and a generated-perturbation analysis."""
import time

import numpy as np
import pytest

from oracle import enkf_oracle as eo
from swfwi_amd import capi, workloads

pytestmark = pytest.mark.gpu

N = 100


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if capi.device_count() < 1:
        pytest.skip("no HIP device")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


def setup(nt):
    case = workloads.ess_301x921(nt=nt)
    vtrue, v0 = workloads.models(case)
    wlt = capi.ricker(case.nt, case.fm, case.dt, case.amp)
    s = capi.Solver(case.nx, case.nz, case.nb, case.dt, case.dx, case.fm, case.nt, case.sx, case.sz, case.gx,
                    case.gz)
    s.bind_velocity(capi.expand_domain(vtrue, case.nb, case.dx, case.dt))
    dobs = workloads.add_noise(s.forward_shots(wlt), 0.01, seed=3)
    members = workloads.ensemble_members(case, v0, N, seed=2)
    return case, s, wlt, dobs, members


def test_enfwi_100_members_match_per_member_oracle(oracle):
    from swfwi_amd.enkf import Enkf, EnsembleFwi
    case, s, wlt, dobs, members = setup(400)
    e = Enkf(s, wlt, dobs, 0.2, N)
    e.set_members(members)
    fwi = EnsembleFwi(e, case.vmin, case.vmax, 5, 200.0)
    print(f"ensemble epoch N={N} on {case.name} nt={case.nt}", flush=True)
    t0 = time.perf_counter()
    fwi.epoch(0)
    print(f"  {1e3 * (time.perf_counter() - t0):.0f} ms", flush=True)
    st = fwi.stats()
    got_vel = e.members()
    got_grad = fwi.gradients()
    assert st["trials"] == sum(int(m["ntrial"]) for m in st["members"])
    g = oracle.Geom(case.nx, case.nz, case.nb, case.dt, case.dx, case.fm, case.nt, case.sx, case.sz, case.gx,
                    case.gz)
    oracle.set_sum_mode(1)  # FP64 accumulation, as the GPU reductions
    try:
        for i in (0, 57, 99):
            print(f"  oracle member {i}", flush=True)
            ref = oracle.Ess(g, members[i].copy(), wlt, dobs, case.vmin, case.vmax, 5, 200.0)
            ref.epoch(0)
            a, b = st["members"][i], ref.scalars()
            for k in ("obj1", "steplen", "alpha2", "alpha3", "obj2", "obj3", "updateobj", "ntrial", "parabolic"):
                assert abs(a[k] - b[k]) <= 1e-6 * max(abs(b[k]), 1e-30), (i, k, a[k], b[k])
            assert rel_l2(got_grad[i], ref.grad()) < 1e-6, i
            assert rel_l2(got_vel[i], ref.vel) < 1e-6, i
    finally:
        oracle.set_sum_mode(0)
    fwi.close()
    e.close()
    s.close()


def test_enkf_100_members_full_size_against_svd(oracle):
    """One batched epoch of all members at nt=4000, then the analysis
    (main-enfwi-damp.cpp:559-605 runs the epochs, then the analysis every
    niterenkf iterations).  The analysis adds the update t5 (transformed
    units) to velocities in m/s as the reference writes it
    (enkfanalyze.cpp:144-147,171-180), so members may leave the scheme's
    stable Courant range; the range is printed, not assumed."""
    from swfwi_amd.enkf import Enkf, EnsembleFwi
    case, s, wlt, dobs, members = setup(4000)
    nd = case.nt * case.ng
    e = Enkf(s, wlt, dobs, 0.2, N)
    e.set_members(members)
    fwi = EnsembleFwi(e, case.vmin, case.vmax, 5, 200.0)
    print(f"ensemble epoch N={N} nt={case.nt}", flush=True)
    t0 = time.perf_counter()
    fwi.epoch(0)
    print(f"  {1e3 * (time.perf_counter() - t0):.0f} ms", flush=True)
    st = fwi.stats()
    after = e.members()
    assert np.all(np.isfinite(after))
    assert np.count_nonzero(np.any(after != members, axis=(1, 2))) == N  # every member moved
    for m in st["members"]:
        assert np.isfinite(m["obj1"]) and m["obj1"] > 0 and np.isfinite(m["steplen"])
        assert m["ntrial"] >= 2
    obj = fwi.mean_objective()
    assert np.isfinite(obj) and obj > 0
    members = after

    # H on each member with the analysis codes (ENKF_SEED=2, first draw),
    # encoded forward modelling without mute: one run per member on the GPU
    # (the forward runs are pinned bit-exactly to the oracle by the parity
    # tests; the oracle would need minutes for 100 runs of 4000 steps)
    code = oracle.gen_codes(2, case.ns)
    enc = np.ascontiguousarray(wlt[:, None] * code[None, :].astype(np.float32))
    d = oracle.encode_obs(code, dobs)
    ha = np.empty((N, case.nt, case.ng), np.float32)
    h = capi.Solver(case.nx, case.nz, case.nb, case.dt, case.dx, case.fm, case.nt, case.sx, case.sz, case.gx,
                    case.gz)
    for i in range(N):
        h.bind_velocity(members[i])
        ha[i] = h.forward_modeling(enc, shot=-1)
    h.close()
    # the perturbation at the reference's scale: N(0, sigma), sigma =
    # sigfac * max|HA'| (initPerturbSigma, enkfanalyze.cpp:497-499,637-654);
    # injected so that numpy sees the same matrix
    hap = ha.reshape(N, -1).astype(np.float64)
    hap -= hap.mean(axis=0, keepdims=True)
    sigma = float(np.abs(hap).max()) * float(np.float32(0.2))  # sigfac is a float
    del hap
    rng = np.random.default_rng(4)
    E = rng.standard_normal((N, case.nt, case.ng)) * sigma
    print(f"enkf analysis N={N} nd={nd} sigma={sigma:.4e}", flush=True)
    e.analyze(E)
    got = e.members()
    info = e.info()
    assert abs(info["sigma"] - sigma) <= 1e-12 * sigma
    print("  numpy SVD of the 3.7M x 100 band", flush=True)
    t0 = time.perf_counter()
    want, mid = eo.analyze(members, ha, d, E, case.dx, case.dt, sigfac=0.2)
    print(f"  {time.perf_counter() - t0:.1f} s; clip {mid['clip']} (GPU {info['clip']}), "
          f"s_1 {mid['svals'][0]:.4e}, s_clip {mid['svals'][mid['clip'] - 1]:.4e}", flush=True)
    del ha, E
    S, Sg = mid["svals"], info["svals"]
    clip = mid["clip"]
    assert info["clip"] == clip and 1 <= clip <= N
    # The GPU forms the Gram matrix band^T band (FP64 sums over 3.7M rows,
    # relative rounding ~ n_rows * 2^-53 of s_1^2 at worst) and takes
    # s_i = sqrt(lambda_i): the error of s_i is bounded by
    # eps_G * s_1^2 / (2 s_i) with eps_G = 1e-10 (rows * eps with margin).
    eps_g = 1e-10
    bound = eps_g * S[0] ** 2 / (2 * S[:clip]) + 1e-12 * S[:clip]
    serr = np.abs(Sg[:clip] - S[:clip])
    print(f"  singular values: max |err|/bound {np.max(serr / bound):.3e}, max rel {np.max(serr / S[:clip]):.3e}",
          flush=True)
    assert np.all(serr <= bound)
    t4 = mid["t4"]
    gerr = np.abs(info["gain"] - t4).max() / np.abs(t4).max()
    print(f"  t4 max rel {gerr:.3e}", flush=True)
    # t4 = P V S^-4 V^T Q: the smallest retained lambda (squared condition of
    # the band) carries the Gram rounding into t4 at eps_G (s_1/s_clip)^2
    assert gerr <= max(1e-6, 4 * eps_g * (S[0] / S[clip - 1]) ** 2)
    rel = np.abs(got.astype(np.float64) - want) / np.abs(want)
    print(f"  members max rel {rel.max():.3e}", flush=True)
    assert rel.max() < 1e-5
    assert np.all(np.isfinite(got)) and np.abs(got - members).max() > 0
    v = case.dx / (case.dt * np.sqrt(got.astype(np.float64)))
    print(f"  analysed velocities {v.min():.1f} .. {v.max():.1f} m/s (Courant "
          f"{v.max() * case.dt / case.dx:.3f} at the fastest point)", flush=True)

    # a generated-perturbation analysis on the same members
    e.set_members(members)
    e.analyze(None, seed=5)
    info2 = e.info()
    assert 1 <= info2["clip"] <= N and info2["sigma"] == info["sigma"]
    assert np.all(np.isfinite(e.members())) and not np.array_equal(members, e.members())
    fwi.close()
    e.close()
    s.close()
