"""GPU parity: the HIP path through the C ABI against the CPU oracle on the
same seeded inputs.

Tolerances: in exact mode the kernels reproduce the reference's expression
trees, so seismograms and gradients must be bit-identical; in FP32 mode the
FP32 budgets apply (seismogram relative L2 <= 1e-5, gradient relative L2 <=
1e-4)."""
import numpy as np
import pytest

from swfwi_amd import capi, workloads

pytestmark = pytest.mark.gpu

SEIS_TOL = 1e-5
GRAD_TOL = 1e-4


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if capi.device_count() < 1:
        pytest.skip("no HIP device")


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def rel_l2(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def setup_case(oracle, case, precision=capi.PREC_EXACT):
    vtrue, v0 = workloads.models(case)
    vt = capi.expand_domain(vtrue, case.nb, case.dx, case.dt)
    vi = capi.expand_domain(v0, case.nb, case.dx, case.dt)
    s = capi.Solver(case.nx, case.nz, case.nb, case.dt, case.dx, case.fm, case.nt, case.sx, case.sz, case.gx,
                    case.gz, precision)
    g = oracle.Geom(case.nx, case.nz, case.nb, case.dt, case.dx, case.fm, case.nt, case.sx, case.sz, case.gx,
                    case.gz)
    wlt = capi.ricker(case.nt, case.fm, case.dt, case.amp)
    return s, g, vt, vi, wlt


# -- per-step kernels ----------------------------------------------------------

@pytest.mark.parametrize("nx,nz,nb", [(48, 40, 14), (353, 233, 26), (101, 77, 36), (19, 13, 4)])
def test_step_kernels_bitexact(oracle, nx, nz, nb):
    print(f"step kernels nx={nx} nz={nz} nb={nb}", flush=True)
    rng = np.random.default_rng(nx * 1000 + nz)
    cur = rng.standard_normal((nx, nz)).astype(np.float32)
    prev = rng.standard_normal((nx, nz)).astype(np.float32)
    vel = rng.uniform(4.9, 44.0, (nx, nz)).astype(np.float32)
    sl = (slice(5, nx - 5), slice(5, nz - 5))
    a, ua = prev.copy(), np.zeros_like(cur)
    capi.load().fd4t10s_damp_zjh_2d_vtrans(a, cur, vel, ua, nx, nz, nb)
    b, ub = prev.copy(), np.zeros_like(cur)
    oracle.fd4t10s_damp(b, cur, vel, ub, nb)
    assert np.array_equal(bits(a), bits(b))
    assert np.array_equal(bits(ua[sl]), bits(ub[sl]))
    a, ua = prev.copy(), np.zeros_like(cur)
    capi.load().fd4t10s_zjh_2d_vtrans(a, cur, vel, ua, nx, nz)
    b, ub = prev.copy(), np.zeros_like(cur)
    oracle.fd4t10s(b, cur, vel, ub)
    assert np.array_equal(bits(a), bits(b))
    assert np.array_equal(bits(ua[sl]), bits(ub[sl]))
    # FP32 mode: same step within a few ulps
    c, uc = prev.copy(), np.zeros_like(cur)
    capi.step_damped(c, cur, vel, uc, nb, capi.PREC_FP32)
    b, ub = prev.copy(), np.zeros_like(cur)
    oracle.fd4t10s_damp(b, cur, vel, ub, nb)
    assert rel_l2(c, b) < 1e-6


def test_step_matches_golden_fixture():
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "fd4t10s_golden.npz"))
    cur, prev, vel, nb = g["c0_cur"], g["c0_prev"], g["c0_vel"], int(g["c0_nb"])
    p, u = prev.copy(), np.zeros_like(cur)
    capi.step_damped(p, cur, vel, u, nb)
    assert np.array_equal(bits(p), bits(g["c0_damp_next"]))


# -- time loops ----------------------------------------------------------------

@pytest.fixture(scope="module")
def small(oracle):
    case = workloads.small_case()
    s, g, vt, vi, wlt = setup_case(oracle, case)
    s.bind_velocity(vt)
    return case, s, g, vt, vi, wlt


def test_boundary_band_pack_unpack(oracle, small):
    """writeBndry / readBndry of one step through the C ABI against the
    oracle's restatement (a random field: every band point distinct)."""
    case, s, g, vt, vi, wlt = small
    rng = np.random.default_rng(17)
    field = rng.standard_normal(s.shape).astype(np.float32)
    b = s.write_bndry(field)
    assert b.size == s.bndr_size
    assert np.array_equal(bits(b), bits(g.write_bndr(field)))
    base = rng.standard_normal(s.shape).astype(np.float32)
    want = base.copy()
    g.read_bndr(b, want)
    got = s.read_bndry(b, base)
    assert np.array_equal(bits(got), bits(want))
    # exactly the band points changed (random values: no coincidences)
    assert np.count_nonzero(got != base) == s.bndr_size


def test_fwi_forward_modeling_bitexact(oracle, small):
    case, s, g, vt, vi, wlt = small
    for shot in (0, case.ns - 1):
        print(f"FwiForwardModeling shot={shot}", flush=True)
        d_gpu = s.forward_modeling(wlt, shot=shot)
        d_cpu = g.forward_modeling(vt, wlt, shot=shot)
        assert np.array_equal(bits(d_gpu), bits(d_cpu))


def test_ess_forward_modeling_bitexact(oracle, small):
    case, s, g, vt, vi, wlt = small
    code = capi.random_codes(1, case.ns)
    enc = oracle.encode_source(code, wlt)
    d_gpu = s.forward_modeling(enc, shot=-1)
    d_cpu = g.forward_modeling(vt, enc, shot=-1)
    assert np.array_equal(bits(d_gpu), bits(d_cpu))


def test_forward_shots_bitexact(oracle, small):
    case, s, g, vt, vi, wlt = small
    dobs = s.forward_shots(wlt)
    for shot in range(case.ns):
        assert np.array_equal(bits(dobs[shot]), bits(g.forward_modeling(vt, wlt, shot=shot)))


def test_tile_tuner_choice_and_results(oracle):
    """The tuner (on by default) picks a width per kernel kind and batch size
    on the first forward run; with it off the measured rule decides (64x8
    tiles for this small grid, fewer than 600 64x16 tiles); the seismograms
    are the oracle's either way, bit for bit."""
    case = workloads.small_case(nx=60, nz=40, nb=8, nt=400, ns=3)
    s, g, vt, vi, wlt = setup_case(oracle, case)
    s.bind_velocity(vt)
    want = np.stack([g.forward_modeling(vt, wlt, shot=k) for k in range(case.ns)])
    for tuning in (True, False):
        print(f"tuning={tuning}", flush=True)
        s.set_tuning(tuning)
        d = s.forward_shots(wlt)
        assert np.array_equal(bits(d), bits(want))
        widths = (s.tile_x(0, case.ns), s.tile_x(1, case.ns))
        print(f"  widths forward/backward at {case.ns} entries: {widths}", flush=True)
        assert all(w in (8, 16, 32) for w in widths)
        if not tuning:
            assert widths == (8, 8)
    s.close()


def test_graph_capture_at_every_batch_size_is_bitexact(oracle):
    """Graph mode 2 captures the 12-shot batch (past the 8-entry limit of
    the default mode) and replays it on the second call: both calls give the
    directly launched seismograms and the oracle's, bit for bit."""
    case = workloads.small_case(nx=60, nz=40, nb=8, nt=400, ns=12)
    s, g, vt, vi, wlt = setup_case(oracle, case)
    s.bind_velocity(vt)
    s.set_graphs(0)
    direct = s.forward_shots(wlt)
    s.set_graphs(2)
    for call in range(2):
        print(f"graph mode 2, call {call}", flush=True)
        assert np.array_equal(bits(s.forward_shots(wlt)), bits(direct))
    for shot in (0, case.ns - 1):
        assert np.array_equal(bits(direct[shot]), bits(g.forward_modeling(vt, wlt, shot=shot)))
    with pytest.raises(capi.SwfwiError):
        s.set_graphs(3)
    s.close()


@pytest.mark.parametrize("shot", [1, -1])
def test_calgradient_bitexact(oracle, small, shot):
    case, s, g, vt, vi, wlt = small
    print(f"calgradient shot={shot}", flush=True)
    rng = np.random.default_rng(11)
    vsrc = rng.standard_normal((case.nt, case.ng)).astype(np.float32)
    enc = wlt if shot >= 0 else oracle.encode_source(capi.random_codes(1, case.ns), wlt)
    g_gpu = s.calgradient(enc, vsrc, shot=shot)
    g_cpu = g.calgradient(vt, enc, vsrc, shot=shot)
    assert np.count_nonzero(g_cpu) > 0
    assert np.array_equal(bits(g_gpu), bits(g_cpu))


@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("tile_x", [8, 16, 32])
def test_every_tile_width_is_bitexact(oracle, small, tile_x, split):
    """Each instantiated tile width, with the backward iteration fused or
    split into source and receiver blocks, reproduces the oracle's seismogram
    and gradient bit for bit."""
    case, s, g, vt, vi, wlt = small
    print(f"tile_x={tile_x} split={split}", flush=True)
    s.set_tile_x(tile_x)
    s.set_bwd_split(split)
    try:
        d_gpu = s.forward_modeling(wlt, shot=0)
        assert s.tile_x(0, 1) == tile_x
        assert np.array_equal(bits(d_gpu), bits(g.forward_modeling(vt, wlt, shot=0)))
        rng = np.random.default_rng(13)
        vsrc = rng.standard_normal((case.nt, case.ng)).astype(np.float32)
        assert np.array_equal(bits(s.calgradient(wlt, vsrc, shot=0)), bits(g.calgradient(vt, wlt, vsrc, shot=0)))
        dobs = s.forward_shots(wlt)
        assert np.array_equal(bits(dobs[case.ns - 1]), bits(g.forward_modeling(vt, wlt, shot=case.ns - 1)))
    finally:
        s.set_tile_x(0)
        s.set_bwd_split(-1)


@pytest.mark.parametrize("nt", [361, 362])
def test_sources_near_tile_edges(oracle, nt):
    """Sources and receivers falling near tile edges, on a grid whose last
    tile row and column are partial: the ESS seismogram and gradient stay
    bit-identical for every tile width."""
    case = workloads.small_case(nx=70, nz=50, nb=8, nt=nt, ns=5)
    s, g, vt, vi, wlt = setup_case(oracle, case)
    s.bind_velocity(vt)
    code = capi.random_codes(3, case.ns)
    enc = oracle.encode_source(code, wlt)
    print(f"tile edges nt={nt}", flush=True)
    rng = np.random.default_rng(nt)
    vsrc = rng.standard_normal((case.nt, case.ng)).astype(np.float32)
    want_d = g.forward_modeling(vt, enc, shot=-1)
    want_g = g.calgradient(vt, enc, vsrc, shot=-1)
    try:
        for tx in (8, 16, 32):
            s.set_tile_x(tx)
            assert np.array_equal(bits(s.forward_modeling(enc, shot=-1)), bits(want_d)), tx
            assert np.array_equal(bits(s.calgradient(enc, vsrc, shot=-1)), bits(want_g)), tx
    finally:
        s.set_tile_x(0)


def test_coincident_sources_accumulate_in_id_order(oracle):
    """Several coded sources on one grid point (and receivers doubled up):
    the injected amplitudes add in source-id order, as the reference's
    sequential loop does, at step 0 and inside the step kernels."""
    from swfwi_amd.workloads import Case
    nx, nz = 40, 30
    sx = np.array([10, 10, 22, 10, 22], np.int32)
    gx = np.concatenate([np.arange(nx), [5, 5]]).astype(np.int32)
    case = Case("coincident", nz, nx, 10.0, 8, 0.001, 340, 25.0, sx, np.ones_like(sx), gx,
                np.zeros(gx.size, np.int32), 1500.0, 3500.0, seed=9)
    s, g, vt, vi, wlt = setup_case(oracle, case)
    s.bind_velocity(vt)
    code = capi.random_codes(4, case.ns)
    enc = oracle.encode_source(code, wlt) * np.float32(1.37)
    enc[:, 3] *= np.float32(0.3)
    assert np.array_equal(bits(s.forward_modeling(enc, shot=-1)), bits(g.forward_modeling(vt, enc, shot=-1)))
    rng = np.random.default_rng(3)
    vsrc = rng.standard_normal((case.nt, case.ng)).astype(np.float32)
    assert np.array_equal(bits(s.calgradient(enc, vsrc, shot=-1)), bits(g.calgradient(vt, enc, vsrc, shot=-1)))
    for shot in (0, 3):
        assert np.array_equal(bits(s.forward_modeling(wlt, shot=shot)), bits(g.forward_modeling(vt, wlt, shot=shot)))


@pytest.mark.parametrize("split", [0, 1])
def test_crowded_tile_falls_back_to_point_lists(oracle, split):
    """A receiver spread on several rows puts more than 64 receivers (and
    coincident sources) into one tile: the kernels take the global point-list
    path there and stay bit-identical, forward and backward."""
    from swfwi_amd.workloads import Case
    nx, nz = 40, 30
    gx = np.repeat(np.arange(2, 30, dtype=np.int32), 10)
    gz = np.tile(np.arange(0, 10, dtype=np.int32), 28)
    sx = np.array([12] * 70 + [20], np.int32)
    case = Case("crowded", nz, nx, 10.0, 8, 0.001, 330, 25.0, sx, np.ones_like(sx), gx, gz, 1500.0, 3500.0, seed=4)
    s, g, vt, vi, wlt = setup_case(oracle, case)
    s.bind_velocity(vt)
    s.set_tile_x(8)
    s.set_bwd_split(split)
    try:
        code = capi.random_codes(6, case.ns)
        enc = oracle.encode_source(code, wlt)
        assert np.array_equal(bits(s.forward_modeling(enc, shot=-1)), bits(g.forward_modeling(vt, enc, shot=-1)))
        rng = np.random.default_rng(8)
        vsrc = rng.standard_normal((case.nt, case.ng)).astype(np.float32)
        assert np.array_equal(bits(s.calgradient(enc, vsrc, shot=-1)), bits(g.calgradient(vt, enc, vsrc, shot=-1)))
    finally:
        s.set_tile_x(0)
        s.set_bwd_split(-1)


@pytest.mark.parametrize("split", [0, 1])
def test_fp32_mode_within_budget(oracle, small, split):
    case, s, g, vt, vi, wlt = small
    s.set_precision(capi.PREC_FP32)
    s.set_bwd_split(split)
    try:
        d_gpu = s.forward_modeling(wlt, shot=1)
        d_cpu = g.forward_modeling(vt, wlt, shot=1)
        e1 = rel_l2(d_gpu, d_cpu)
        rng = np.random.default_rng(12)
        vsrc = rng.standard_normal((case.nt, case.ng)).astype(np.float32)
        gg = s.calgradient(wlt, vsrc, shot=1)
        gc = g.calgradient(vt, wlt, vsrc, shot=1)
        sl = (slice(s.bx0, s.nxpad - s.bx0), slice(s.bz0, s.nzpad - s.bx0))
        e2 = rel_l2(gg[sl], gc[sl])
    finally:
        s.set_precision(capi.PREC_EXACT)
        s.set_bwd_split(-1)
    print(f"fp32 split={split} seismogram relL2={e1:.3e} gradient relL2={e2:.3e}", flush=True)
    assert e1 <= SEIS_TOL
    assert e2 <= GRAD_TOL


# -- drivers -------------------------------------------------------------------

def test_ess_epochs_match_oracle(oracle, small):
    case, s, g, vt, vi, wlt = small
    dobs = s.forward_shots(wlt)
    oracle.set_sum_mode(1)  # FP64 accumulation, as the GPU reductions
    try:
        s.bind_velocity(vi)
        inv = capi.Inversion(s, wlt, dobs, 1400.0, 3600.0, 3, 200.0)
        e = oracle.Ess(g, vi.copy(), wlt, dobs, 1400.0, 3600.0, 3, 200.0)
        for it in range(2):
            print(f"ess epoch {it}", flush=True)
            inv.ess_epoch(it)
            e.epoch(it)
            a, b = inv.stats(), e.scalars()
            for k in ("obj1", "steplen", "alpha2", "alpha3", "obj2", "obj3", "updateobj", "ntrial", "parabolic"):
                assert abs(a[k] - b[k]) <= 1e-6 * max(abs(b[k]), 1e-30), (it, k, a[k], b[k])
            assert rel_l2(inv.gradient(), e.grad()) < 1e-6
            assert rel_l2(s.velocity(), e.vel) < 1e-6
        inv.close()
    finally:
        oracle.set_sum_mode(0)
        s.bind_velocity(vt)


def test_fwi_epoch_matches_oracle(oracle, small):
    case, s, g, vt, vi, wlt = small
    dobs = s.forward_shots(wlt)
    oracle.set_sum_mode(1)
    try:
        s.bind_velocity(vi)
        inv = capi.Inversion(s, wlt, dobs, 1400.0, 3600.0, 3, 200.0)
        e = oracle.Ess(g, vi.copy(), wlt, dobs, 1400.0, 3600.0, 3, 200.0)
        for it in range(2):
            print(f"fwi epoch {it}", flush=True)
            inv.fwi_epoch(it)
            e.fwi_epoch(it)
            a, b = inv.stats(), e.scalars()
            for k in ("obj1", "steplen", "obj2", "obj3", "updateobj"):
                assert abs(a[k] - b[k]) <= 1e-6 * max(abs(b[k]), 1e-30), (it, k, a[k], b[k])
            assert np.array_equal(bits(inv.gradient()), bits(e.grad()))
            assert rel_l2(s.velocity(), e.vel) < 1e-6
        inv.close()
    finally:
        oracle.set_sum_mode(0)
        s.bind_velocity(vt)


def test_fwi_201x301_shot_full_size(oracle):
    """One shot of the 201x301 workload at full size (nb=20, nt=2000):
    seismogram and raw gradient bit-identical to the oracle."""
    case = workloads.fwi_201x301()
    s, g, vt, vi, wlt = setup_case(oracle, case)
    s.bind_velocity(vi)
    print("fwi_201x301 forward", flush=True)
    d_gpu = s.forward_modeling(wlt, shot=5)
    d_cpu = g.forward_modeling(vi, wlt, shot=5)
    assert np.array_equal(bits(d_gpu), bits(d_cpu))
    rng = np.random.default_rng(5)
    vsrc = (rng.standard_normal((case.nt, case.ng)) * 1e-3).astype(np.float32)
    print("fwi_201x301 gradient", flush=True)
    g_gpu = s.calgradient(wlt, vsrc, shot=5)
    g_cpu = g.calgradient(vi, wlt, vsrc, shot=5)
    assert np.array_equal(bits(g_gpu), bits(g_cpu))


def test_ess_301x921_batched_shots_and_epoch(oracle):
    """The 301x921 grid (nb=30, 60 shots, 921 receivers) at full width, nt cut
    to 800 so the oracle finishes in seconds: the 60-shot batched forward
    (the 64x32 tile path) bit-identical for the first and last shot, and one
    ESS epoch (E=1 forward with band save, split backward, E=2 trials) against
    the FP64-sum oracle."""
    case = workloads.ess_301x921(nt=800)
    s, g, vt, vi, wlt = setup_case(oracle, case)
    s.bind_velocity(vt)
    print("ess_301x921 shots", flush=True)
    dobs = s.forward_shots(wlt)
    for shot in (0, case.ns - 1):
        assert np.array_equal(bits(dobs[shot]), bits(g.forward_modeling(vt, wlt, shot=shot))), shot
    oracle.set_sum_mode(1)
    try:
        s.bind_velocity(vi)
        inv = capi.Inversion(s, wlt, dobs, case.vmin, case.vmax, 5, 200.0)
        e = oracle.Ess(g, vi.copy(), wlt, dobs, case.vmin, case.vmax, 5, 200.0)
        print("ess_301x921 epoch", flush=True)
        inv.ess_epoch(0)
        e.epoch(0)
        a, b = inv.stats(), e.scalars()
        for k in ("obj1", "steplen", "alpha2", "alpha3", "obj2", "obj3", "updateobj", "ntrial", "parabolic"):
            assert abs(a[k] - b[k]) <= 1e-6 * max(abs(b[k]), 1e-30), (k, a[k], b[k])
        assert rel_l2(inv.gradient(), e.grad()) < 1e-6
        assert rel_l2(s.velocity(), e.vel) < 1e-6
        inv.close()
    finally:
        oracle.set_sum_mode(0)


def test_timing_survives_many_epochs(oracle, small):
    """Timing on across 20 ESS epochs: each epoch queues four timed loops (the
    gradient's forward and backward, two line-search trial loops on lane
    streams), so the pending list passes its flush size of 64 while a trial
    lane may still be running; every loop's stop event must be waited on
    before its time is read.  Regression for 'hipEventElapsedTime: device not
    ready' at the 16th timed epoch."""
    case, s, g, vt, vi, wlt = small
    dobs = s.forward_shots(wlt)
    try:
        s.bind_velocity(vi)
        inv = capi.Inversion(s, wlt, dobs, 1400.0, 3600.0, 3, 200.0)
        inv.ess_epoch(0)  # captures the loops into graphs
        s.set_timing(True)
        n0 = s.counters()[1]
        for it in range(1, 21):
            inv.ess_epoch(it)
        kt = s.kernel_times()
        n1 = s.counters()[1]
        s.set_timing(False)
        fwd_ms, fwd_n = kt["fwd"]
        bwd_ms, bwd_n = kt["bwd"]
        assert fwd_ms > 0 and bwd_ms > 0
        assert fwd_n == n1[0] - n0[0] and bwd_n == n1[1] - n0[1], (kt, n0, n1)
        inv.close()
    finally:
        s.bind_velocity(vt)
