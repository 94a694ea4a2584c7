"""Mixed-size batched gradients with reduced-precision checkpoints
(swfwi_batch) through the C ABI.

Groups of different grid sizes and time-step counts run in one launch per
step.  Per entry (one shot of one model):
* the seismogram against the oracle's FwiForwardModeling
  (damp4t10d.cpp:272-289) within the FP32 budget, rel-L2 <= 1e-5 (the batch
  stencil computes in FP32 with fused multiply-adds);
* the adjoint source transVsrc(dobs - dcal) against the oracle's
  (fwiframework.cpp:75-105) within 1e-6;
* the gradient, formed from FP32 / BF16 / FP16 checkpoints every 4 steps,
  against the FP64 run of the reference calgradient (band reconstruction,
  imaging on every step, fwiframework.cpp:114-177), over the region
  maskGradient keeps: rel-L2 <= 1e-3;
* per-model gradients equal the per-entry gradients summed in shot order."""
import numpy as np
import pytest

from swfwi_amd import capi, workloads
from swfwi_amd.batch import Batch, Group

pytestmark = pytest.mark.gpu

SEIS_TOL = 1e-5
VSRC_TOL = 1e-6
GRAD_TOL = 1e-3


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if capi.device_count() < 1:
        pytest.skip("no HIP device")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def groups():
    # two grid sizes, different nt (the shorter group leaves the launches
    # first), two models x three / two shots each
    a = Group(nx=48, nz=32, nb=8, dx=10.0, dt=0.001, fm=12.0, nt=600, sx=np.array([3, 24, 44]),
              sz=np.array([1, 1, 2]), nmodels=2)
    b = Group(nx=70, nz=44, nb=10, dx=10.0, dt=0.001, fm=12.0, nt=720, sx=np.array([10, 55]), sz=np.array([1, 3]),
              nmodels=2)
    return [a, b]


def models_for(g, k):
    case = workloads.Case("b", g.nz, g.nx, g.dx, g.nb, g.dt, g.nt, g.fm, np.asarray(g.sx, np.int32),
                          np.asarray(g.sz, np.int32), np.arange(g.nx, dtype=np.int32), np.zeros(g.nx, np.int32),
                          1500.0, 3500.0, seed=10 + k)
    vtrue, v0 = workloads.models(case)
    return (capi.expand_domain(vtrue, g.nb, g.dx, g.dt), capi.expand_domain(v0, g.nb, g.dx, g.dt))


def inner_mask(g):
    bx0 = g.nb + 6
    m = np.zeros(g.shape, bool)
    m[bx0:g.nxpad - bx0, 6:g.nzpad - bx0] = True
    return m


@pytest.fixture(scope="module")
def run(oracle):
    gs = groups()
    b = Batch(gs, snap_prec=capi.SNAP_F32)
    mods = {}
    for gi, g in enumerate(gs):
        for m in range(g.nmodels):
            mods[(gi, m)] = models_for(g, 2 * gi + m)
    # observed data: the true models
    for (gi, m), (vt, vi) in mods.items():
        b.bind_model(gi, m, vt)
    print("batch forward on the true models", flush=True)
    dobs = b.forward()
    # the initial models: forward with checkpoints, adjoint sources
    for (gi, m), (vt, vi) in mods.items():
        b.bind_model(gi, m, vi)
    print("batch forward on the initial models", flush=True)
    dcal = b.forward()
    vsrc = b.adjoint_source(dobs)
    return gs, b, mods, dobs, dcal, vsrc


def oracle_geom(oracle, g):
    return oracle.Geom(g.nx, g.nz, g.nb, g.dt, g.dx, g.fm, g.nt, np.asarray(g.sx, np.int32),
                       np.asarray(g.sz, np.int32), np.arange(g.nx, dtype=np.int32), np.zeros(g.nx, np.int32))


def test_batch_seismograms_and_adjoint_sources(oracle, run):
    gs, b, mods, dobs, dcal, vsrc = run
    D, V, O = b.split_data(dcal), b.split_data(vsrc), b.split_data(dobs)
    for k, (gi, m, s) in enumerate(b.entries):
        g = gs[gi]
        og = oracle_geom(oracle, g)
        wlt = oracle.ricker(g.nt, g.fm, g.dt, 1.0)
        want = og.forward_modeling(mods[(gi, m)][1], wlt, shot=s)
        e = rel_l2(D[k], want)
        assert e <= SEIS_TOL, (gi, m, s, e)
        r = (O[k] - D[k]).astype(np.float32)
        oracle.trans_vsrc(r)
        assert rel_l2(V[k], r) <= VSRC_TOL, (gi, m, s)


@pytest.mark.parametrize("prec", [capi.SNAP_F32, capi.SNAP_BF16, capi.SNAP_F16])
def test_batch_gradient_against_fp64_oracle(oracle, run, prec):
    gs, b, mods, dobs, dcal, vsrc = run
    if prec != capi.SNAP_F32:
        # a fresh batch in the narrow format (its checkpoints are half the size)
        b = Batch(gs, snap_prec=prec)
        for (gi, m), (vt, vi) in mods.items():
            b.bind_model(gi, m, vi)
        b.forward(want_data=False)
    print(f"batch backward, checkpoints {['f32', 'bf16', 'f16'][prec]}", flush=True)
    ge, gm = b.backward(vsrc)
    GE, GM = b.split_grid(ge), b.split_grid(gm, per_model=True)
    V = b.split_data(vsrc)
    errs = []
    for k, (gi, m, s) in enumerate(b.entries):
        g = gs[gi]
        og = oracle_geom(oracle, g)
        wlt = oracle.ricker(g.nt, g.fm, g.dt, 1.0)
        want = og.calgradient64(mods[(gi, m)][1], wlt, V[k], shot=s)
        mask = inner_mask(g)
        e = rel_l2(GE[k][mask], want[mask])
        errs.append(e)
        assert e <= GRAD_TOL, (gi, m, s, e)
    print(f"  gradient rel-L2 vs FP64: max {max(errs):.2e}, mean {np.mean(errs):.2e}", flush=True)
    # model gradients: the shots summed in shot order
    mi = 0
    for gi, g in enumerate(gs):
        for m in range(g.nmodels):
            ks = [k for k, (a, bm, s) in enumerate(b.entries) if a == gi and bm == m]
            acc = np.zeros(g.shape, np.float32)
            for k in ks:
                acc = (acc + GE[k]).astype(np.float32)
            assert np.array_equal(GM[mi], acc)
            mi += 1
    info = b.info()
    assert info["entries"] == len(b.entries) and info["snapshot_bytes"] > 0
    if prec != capi.SNAP_F32:
        b.close()


def test_wide_tile_batch_against_oracles(oracle):
    """A batch big enough for the wide step tiles (24 entries of the 301x921
    grid: 8928 tiles of 64 x 16, above the 8000 from which the batch's rule
    takes 24-wide tiles), run with 32-wide tiles, nt=800: seismograms of the first and last
    entry against the FP32 oracle, their gradients from FP32 checkpoints
    against the FP64 oracle run."""
    g = Group(nx=921, nz=301, nb=30, dx=12.5, dt=0.001, fm=10.0, nt=800,
              sx=np.linspace(0, 920, 12).astype(np.int32), sz=np.ones(12, np.int32), nmodels=2)
    ux, uz = g.nx + 2 * g.nb, g.nz + g.nb
    assert g.nmodels * len(g.sx) * ((uz + 63) // 64) * ((ux + 15) // 16) >= 8000
    b = Batch([g], snap_prec=capi.SNAP_F32)
    b.set_tile_x(32)  # fixed: the widest tile, which neither the rule nor the tuner need pick
    mods = [models_for(g, 7 + m) for m in range(2)]
    for m in range(2):
        b.bind_model(0, m, mods[m][0])
    dobs = b.forward()
    assert b.tile_x() == 32
    for m in range(2):
        b.bind_model(0, m, mods[m][1])
    D = b.split_data(b.forward())
    V = b.split_data(b.adjoint_source(dobs))
    GE = b.split_grid(b.backward(None, want_entries=True, want_models=False)[0])
    og = oracle_geom(oracle, g)
    wlt = oracle.ricker(g.nt, g.fm, g.dt, 1.0)
    mask = inner_mask(g)
    for k in (0, len(b.entries) - 1):
        _, m, s = b.entries[k]
        want = og.forward_modeling(mods[m][1], wlt, shot=s)
        e = rel_l2(D[k], want)
        print(f"  entry {k} (model {m}, shot {s}): seismogram rel-L2 {e:.2e}", flush=True)
        assert e <= SEIS_TOL
        gw = og.calgradient64(mods[m][1], wlt, V[k], shot=s)
        eg = rel_l2(GE[k][mask], gw[mask])
        print(f"  entry {k}: gradient rel-L2 vs FP64 {eg:.2e}", flush=True)
        assert eg <= GRAD_TOL
    b.close()


def test_batch_tile_widths_block_orders_and_tuner():
    """The step kernels at tile widths 16, 24 and 32, with the blocks of a launch
    entry by entry (order 0) or a model's shots tile by tile (order 1), give
    the mixed-size batch's seismograms and gradients (BF16 checkpoints) bit
    for bit alike: each output point is formed by the same operations in the
    same order whatever the tile and the block order.  With the tuner (the
    default) the first forward run picks one of them; with it off the rule
    decides (width 16 for this small batch, order 1)."""
    gs = groups()
    mods = {(gi, m): models_for(g, 2 * gi + m) for gi, g in enumerate(gs) for m in range(g.nmodels)}
    out = {}
    modes = [(16, 0), (16, 1), (24, 1), (32, 0), (32, 1), "tuner", "rule"]
    for mode in modes:
        b = Batch(gs, snap_prec=capi.SNAP_BF16)
        if mode == "rule":
            b.set_tuning(False)
        elif mode != "tuner":
            b.set_tile_x(mode[0])
            b.set_block_order(mode[1])
        for (gi, m), (vt, vi) in mods.items():
            b.bind_model(gi, m, vi)
        d = b.forward()
        got = (b.tile_x(), b.block_order())
        print(f"{mode}: tile width {got[0]}, block order {got[1]}", flush=True)
        assert got[0] in (16, 24, 32) and got[1] in (0, 1)
        if isinstance(mode, tuple):
            assert got == mode
        if mode == "rule":
            assert got == (16, 1)
        vsrc = (np.random.default_rng(5).standard_normal(d.size) * 1e-3).astype(np.float32)
        ge, gm = b.backward(vsrc)
        assert np.count_nonzero(ge) > 0
        out[mode] = (d, ge, gm)
        b.close()
    for mode in modes[1:]:
        for k in range(3):
            assert np.array_equal(out[mode][k].view(np.uint32), out[(16, 0)][k].view(np.uint32)), (mode, k)
