"""EnKF analysis: the numpy oracle pinned by linear algebra (CPU), and the GPU
analysis against the oracle on injected perturbations (GPU)."""
import numpy as np
import pytest

from oracle import enkf_oracle as eo


def test_clip_position_against_c_oracle_and_hand_case(oracle):
    rng = np.random.default_rng(0)
    for _ in range(20):
        s = np.sort(rng.lognormal(0, 3, 12))[::-1].copy()
        assert eo.clip_position(s) == oracle.clip_position(s)
    # hand case: rel/(1-rel)*10 = 0.01001 > 4*0.001 only at pos 1 -> clip 2
    s = np.array([10.0, 1.0, 0.001, 1e-9])
    assert eo.clip_position(s) == oracle.clip_position(s) == 2


def test_gram_eigen_form_equals_svd_form():
    """t4 = HA'^T U_k S_k^-2 U_k^T t0 (the reference's SVD form) equals
    P V_k S_k^-4 V_k^T Q with P = HA'^T band, Q = band^T t0 and band^T band =
    V S^2 V^T (the form the GPU computes)."""
    rng = np.random.default_rng(1)
    nd, N = 500, 7
    hap = rng.standard_normal((nd, N))
    hap -= hap.mean(axis=1, keepdims=True)
    band = hap + 0.1 * rng.standard_normal((nd, N))
    t0 = rng.standard_normal((nd, N))
    U, S, _ = np.linalg.svd(band, full_matrices=False)
    k = 5
    t4_svd = hap.T @ U[:, :k] @ np.diag(1 / S[:k] ** 2) @ U[:, :k].T @ t0
    lam, V = np.linalg.eigh(band.T @ band)
    order = np.argsort(lam)[::-1]
    lam, V = lam[order], V[:, order]
    P, Q = hap.T @ band, band.T @ t0
    t4_gram = P @ V[:, :k] @ np.diag(1 / lam[:k] ** 2) @ V[:, :k].T @ Q
    assert np.allclose(np.sqrt(lam), S, rtol=1e-10)
    assert np.allclose(t4_gram, t4_svd, rtol=1e-8, atol=1e-10 * np.abs(t4_svd).max())


def test_member_mean_is_fp32_member_order():
    rng = np.random.default_rng(2)
    m = rng.uniform(5, 40, (9, 7, 5)).astype(np.float32)
    mean = eo.member_mean(m)
    acc = np.float32(0)
    for j in range(9):
        acc = np.float32(acc + m[j, 3, 2])
    assert mean[3, 2] == np.float32(acc / np.float32(9))


# -- GPU ----------------------------------------------------------------------

def _ensemble(case, N, seed):
    from swfwi_amd import capi, workloads
    vtrue, v0 = workloads.models(case)
    rng = np.random.default_rng(seed)
    mem = []
    from scipy.ndimage import gaussian_filter
    for i in range(N):
        f = gaussian_filter(rng.standard_normal(v0.shape), 3.0)
        f = f / np.abs(f).max() * 0.03
        mem.append(capi.expand_domain((v0 * (1 + f)).astype(np.float32), case.nb, case.dx, case.dt))
    return vtrue, np.stack(mem)


@pytest.mark.gpu
def test_gpu_enkf_analysis_matches_oracle(oracle):
    from swfwi_amd import capi, workloads
    from swfwi_amd.enkf import Enkf
    if capi.device_count() < 1:
        pytest.skip("no HIP device")
    case = workloads.small_case(nx=48, nz=32, nb=8, nt=360, ns=3)
    N = 6
    vtrue, members = _ensemble(case, N, 3)
    wlt = capi.ricker(case.nt, case.fm, case.dt, 1.0)
    s = capi.Solver(case.nx, case.nz, case.nb, case.dt, case.dx, case.fm, case.nt, case.sx, case.sz, case.gx,
                    case.gz)
    s.bind_velocity(capi.expand_domain(vtrue, case.nb, case.dx, case.dt))
    dobs = s.forward_shots(wlt)
    e = Enkf(s, wlt, dobs, 0.2, N)
    e.set_members(members)
    assert np.array_equal(e.mean(), eo.member_mean(members))
    rng = np.random.default_rng(4)
    rms = float(np.sqrt(np.mean(dobs.astype(np.float64) ** 2)))
    E = rng.standard_normal((N, case.nt, case.ng)) * 0.01 * rms
    print("enkf analysis N=6", flush=True)
    e.analyze(E)
    got = e.members()
    info = e.info()
    # oracle: H on each member with the analysis codes (ENKF_SEED=2), no mute
    code = oracle.gen_codes(2, case.ns)
    g = oracle.Geom(case.nx, case.nz, case.nb, case.dt, case.dx, case.fm, case.nt, case.sx, case.sz, case.gx,
                    case.gz)
    enc = oracle.encode_source(code, wlt)
    d = oracle.encode_obs(code, dobs)
    ha = np.stack([g.forward_modeling(members[i], enc, shot=-1) for i in range(N)])
    want, mid = eo.analyze(members, ha, d, E, case.dx, case.dt, sigfac=0.2)
    assert info["clip"] == mid["clip"]
    S, Sg, clip = mid["svals"], info["svals"], mid["clip"]
    assert np.allclose(Sg[:clip], S[:clip], rtol=1e-7, atol=1e-9 * S[0])
    # past clip (discarded; here the null direction of the row-centred band,
    # zero in exact arithmetic): the Gram form band^T band = V S^2 V^T
    # resolves s only to sqrt(eps_G) s_1, eps_G = 1e-10 the Gram rounding
    assert np.all(np.abs(Sg[clip:] - S[clip:]) <= np.sqrt(1e-10) * S[0])
    assert np.allclose(info["sigma"], mid["sigma"], rtol=1e-12)
    t4 = mid["t4"]
    assert np.abs(info["gain"] - t4).max() <= 1e-6 * np.abs(t4).max()
    rel = np.abs(got.astype(np.float64) - want) / np.abs(want)
    assert rel.max() < 1e-5, rel.max()
    # the analysis moved the members (the gain is not trivially zero)
    assert np.abs(got - members).max() > 0


@pytest.mark.gpu
def test_gpu_enkf_generated_perturbation_runs_twice():
    from swfwi_amd import capi, workloads
    from swfwi_amd.enkf import Enkf
    if capi.device_count() < 1:
        pytest.skip("no HIP device")
    case = workloads.small_case(nx=40, nz=30, nb=8, nt=330, ns=2)
    N = 5
    vtrue, members = _ensemble(case, N, 7)
    wlt = capi.ricker(case.nt, case.fm, case.dt, 1.0)
    s = capi.Solver(case.nx, case.nz, case.nb, case.dt, case.dx, case.fm, case.nt, case.sx, case.sz, case.gx,
                    case.gz)
    s.bind_velocity(capi.expand_domain(vtrue, case.nb, case.dx, case.dt))
    dobs = s.forward_shots(wlt)
    e = Enkf(s, wlt, dobs, 0.2, N)
    e.set_members(members)
    e.analyze(None, seed=11)
    sig1 = e.info()["sigma"]
    a1 = e.members()
    e.analyze(None, seed=12)
    info = e.info()
    assert sig1 > 0 and info["sigma"] == sig1  # fixed at the first analysis
    assert 1 <= info["clip"] <= N
    assert np.all(np.isfinite(e.members())) and not np.array_equal(a1, e.members())
