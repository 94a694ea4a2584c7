"""Benchmark: encoded-source (ESS) FWI on the 301x921 grid (nb=30, nt=4000,
60 Ricker 10 Hz shots, 921 receivers), one GPU.

A step is one EssFwiFramework epoch on the GPU: new +-1 codes, encoded
forward modelling + boundary save (one fused pass), mute, residual, adjoint
gradient, PR+ direction, bracketing/parabolic line search (the trial
forward runs side by side on their own streams), model update.  value = stencil grid-point updates per second
over the timed epochs.

    python bench.py [--steps K] [--warmup W] [--precision exact|fp32]
    python bench.py --workload fwi_201x301   # plain FWI, 201x301, 11 shots:
                                             # FwiFramework epochs, gradient per shot
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FWD_BYTES_PER_PT = 16.0   # read cur, prev, model; write next
BWD_BYTES_PER_PT = 40.0   # two steps (source + receiver field) + gradient read-modify-write


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` in this workload, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs summarised under profiles/ (null if absent)."""
    path = os.path.join(ROOT, "profiles", f"k_fwd_step_{workload}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch:" and f[1] == kernel:
            return int(f[2])
    return None


def cpu_epoch(case, encoded, dobs):
    """One whole epoch of the oracle with the reference's threading (ESS:
    EssFwiFramework::epoch, essfwiframework.cpp:191-239; plain FWI:
    FwiFramework::epoch over every shot): codes, encoded data, gradient,
    direction and the line search's trial forward runs
    (updatesteplenop.cpp:216-265), on the observed data the GPU computed
    (bit-identical to the oracle's).  Returns (seconds, trial forward runs)."""
    from oracle import pyoracle as po
    from swfwi_amd import workloads

    _, v0 = workloads.models(case)
    g = po.Geom(case.nx, case.nz, case.nb, case.dt, case.dx, case.fm, case.nt, case.sx, case.sz, case.gx, case.gz,
                parallel_u2=0)
    wlt = po.ricker(case.nt, case.fm, case.dt, case.amp)
    inv = po.Ess(g, po.expand_domain(v0, case.nb, case.dx, case.dt), wlt, dobs, case.vmin, case.vmax, 5, 200.0)
    t0 = time.perf_counter()
    (inv.epoch if encoded else inv.fwi_epoch)(0)
    return time.perf_counter() - t0, inv.scalars()["ntrial"]


def cpu_baseline(case, encoded, seconds_cap=30.0):
    """The oracle (CPU restatement of the reference path, bit-identical to it)
    timed on this host on a bounded sample of the same workload: one gradient
    (ESS: the encoded supershot of the 301x921 grid, nt=4000: 4000 forward
    steps, 3700 backward iterations of two steps, ~11 s with 16 threads;
    plain FWI: one shot of the 201x301 grid, nt=2000), reference threading
    (serial u2 pass of the damped kernel) and all-parallel."""
    from oracle import pyoracle as po
    from swfwi_amd import workloads

    nt = case.nt
    vtrue, v0 = workloads.models(case)
    out = {}
    for label, par in (("reference_threading", 0), ("all_parallel", 1)):
        g = po.Geom(case.nx, case.nz, case.nb, case.dt, case.dx, case.fm, nt, case.sx, case.sz, case.gx, case.gz,
                    parallel_u2=par)
        vi = po.expand_domain(v0, case.nb, case.dx, case.dt)
        wlt = po.ricker(nt, case.fm, case.dt, case.amp)
        src = po.encode_source(po.gen_codes(1, case.ns), wlt) if encoded else wlt
        rng = np.random.default_rng(0)
        vsrc = (rng.standard_normal((nt, case.ng)) * 1e-3).astype(np.float32)
        t0 = time.perf_counter()
        g.calgradient(vi, src, vsrc, shot=-1 if encoded else 0)
        dt = time.perf_counter() - t0
        # the reference's taper test: float dt*it against the double 0.3
        iters = sum(1 for it in range(nt - 1, -1, -1) if float(np.float32(case.dt) * np.float32(it)) > 0.3) + 1
        pts = case.update_points * (nt + 2 * iters)
        out[label] = (pts / dt / 1e9, dt, nt, iters)
        if dt > seconds_cap:
            break
    return out, po.lib().orc_max_threads()


def mixed_groups(shots=16):
    """The mixed-size batch: 4 models of 201x301 (dx=10 m, nb=20, nt=2000,
    15 Hz), 2 of 301x921 (12.5 m, nb=30, nt=4000, 10 Hz) and 1 of 501x1501
    (12.5 m, nb=30, nt=6000, 10 Hz), `shots` Ricker shots each on the surface
    (depth 1 cell), receivers on every surface cell, dt=1 ms."""
    from swfwi_amd.batch import Group

    spec = [(301, 201, 20, 10.0, 15.0, 2000, 4, 1500.0, 4500.0), (921, 301, 30, 12.5, 10.0, 4000, 2, 1500.0, 5500.0),
            (1501, 501, 30, 12.5, 10.0, 6000, 1, 1500.0, 5500.0)]
    groups = []
    for nx, nz, nb, dx, fm, nt, nm, vmin, vmax in spec:
        sx = np.linspace(0, nx - 1, shots).astype(np.int32)
        groups.append(Group(nx=nx, nz=nz, nb=nb, dx=dx, dt=0.001, fm=fm, nt=nt, sx=sx, sz=np.ones(shots, np.int32),
                            gx0=0, ng=nx, gz=0, nmodels=nm, extra=dict(vmin=vmin, vmax=vmax)))
    return groups


def mixed_workload(args):
    """Mixed-size batched gradients with reduced-precision checkpoints: every
    shot of every model in one launch per time step, the source field kept
    every 4 steps in BF16 (value and roofline), FP16 and FP32 (whose error
    against the FP64 oracle run is the imaging subsampling alone); gradients
    of one shot per grid size against the FP64 oracle run."""
    from oracle import pyoracle as po
    from swfwi_amd import capi, workloads
    from swfwi_amd.batch import Batch

    groups = mixed_groups(args.shots)
    models = []
    for gi, g in enumerate(groups):
        for m in range(g.nmodels):
            vt = workloads.layered_model(g.nz, g.nx, g.extra["vmin"], g.extra["vmax"], seed=100 + 10 * gi + m)
            v0 = workloads.smoothed(vt, 20.0)
            models.append((gi, m, capi.expand_domain(vt, g.nb, g.dx, g.dt), capi.expand_domain(v0, g.nb, g.dx, g.dt)))
    res = {}
    timing = {}
    for prec in (capi.SNAP_F32, capi.SNAP_BF16, capi.SNAP_F16):  # FP32 (136 GB of checkpoints) for the baseline
        name = {capi.SNAP_F32: "f32", capi.SNAP_BF16: "bf16", capi.SNAP_F16: "f16"}[prec]
        b = Batch(groups, snap_prec=prec)
        b.set_graphs(not args.no_graphs)
        for gi, m, vt, _ in models:
            b.bind_model(gi, m, vt)
        dobs = b.forward()
        for gi, m, _, vi in models:
            b.bind_model(gi, m, vi)
        b.forward(want_data=False)  # captures the loop (graph) on first use
        b.adjoint_source(dobs, want=False)
        b.backward(None, want_entries=False, want_models=False)
        steps, fms, bms = max(args.steps, 1), [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            b.forward(want_data=False)
            b.adjoint_source(dobs, want=False)
            b.backward(None, want_entries=False, want_models=False)
            inf = b.info()
            fms.append(inf["forward_ms"])
            bms.append(inf["backward_ms"])
        wall = (time.perf_counter() - t0) / steps
        inf = b.info()
        vsrc = b.adjoint_source(dobs)
        ge, _ = b.backward(None, want_entries=True, want_models=False)
        GE, V = b.split_grid(ge), b.split_data(vsrc)
        errs = {}
        for gi, g in enumerate(groups):  # model 0, shot 0 of each size
            k = b.entries.index((gi, 0, 0))
            og = po.Geom(g.nx, g.nz, g.nb, g.dt, g.dx, g.fm, g.nt, g.sx, g.sz, np.arange(g.nx, dtype=np.int32),
                         np.zeros(g.nx, np.int32))
            wlt = po.ricker(g.nt, g.fm, g.dt, 1.0)
            vi = [v for gg, mm, _, v in models if gg == gi and mm == 0][0]
            t1 = time.perf_counter()
            want = og.calgradient64(vi, wlt, V[k], shot=0)
            bx0 = g.nb + 6
            sl = (slice(bx0, g.nxpad - bx0), slice(6, g.nzpad - bx0))
            err = float(np.linalg.norm(GE[k][sl] - want[sl]) / np.linalg.norm(want[sl]))
            errs[f"{g.nz}x{g.nx}"] = round(err, 7)
            print(f"[bench] {name}: gradient of {g.nz}x{g.nx} shot 0 against the FP64 run: rel-L2 {err:.3e} "
                  f"(oracle {time.perf_counter() - t1:.1f} s)", file=sys.stderr, flush=True)
        res[name] = errs
        pts = inf["forward_points"] + 2 * inf["backward_points"]  # the backward advances two fields
        dev_ms = float(np.mean(fms) + np.mean(bms))
        # algorithmic bytes of the step kernels: forward 16 B per point
        # update (read cur, prev, model; write next) + the checkpoints
        # written; backward 16 B per point (the receiver field alone) + the
        # checkpoints read + the gradient read-modify-write on every 4th step
        byts = (16.0 * inf["forward_points"] + inf["snapshot_bytes"] + 16.0 * inf["backward_points"] +
                inf["snapshot_bytes"] + 2.0 * inf["backward_points"])
        timing[name] = {"wall_ms": 1e3 * wall, "device_ms": dev_ms, "forward_ms": float(np.mean(fms)),
                        "backward_ms": float(np.mean(bms)), "points": pts, "snapshot_bytes": inf["snapshot_bytes"],
                        "entries": inf["entries"], "gpts": pts / wall / 1e9, "bytes": byts,
                        "tile_x": b.tile_x(), "block_order": b.block_order(),
                        # one launch per step of the longest entry, forward, and per backward
                        # iteration (the taper ends the backward ~0.3 s early)
                        "launches": max(g.nt for g in groups) + max(g.nt - int(round(0.3 / g.dt)) for g in groups)}
        b.close()
    by_format = {k: {"gpts_per_s": round(v["gpts"], 2), "gradient_ms_per_shot": round(v["device_ms"] / v["entries"], 4),
                     "tile_x": v["tile_x"], "block_order": v["block_order"]} for k, v in timing.items()}
    timing = timing["bf16"]
    gpts = timing["gpts"]
    line = {
        "metric": "FWI gradient time/shot & Gpts/s, learned vs plain FP32 at equal data misfit",
        "value": round(gpts, 3), "unit": "Gpts/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": round(timing["wall_ms"], 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded layered models; observed data from the same propagator on the true models)",
        "config": {"workload": "mixed-size batch: 4x(201x301, nt=2000), 2x(301x921, nt=4000), 1x(501x1501, "
                               f"nt=6000), {args.shots} shots each, one launch per time step, checkpoints every 4 "
                               "steps in BF16 with per-snapshot max-abs scale",
                   "parallelism": "single GPU"},
        "gradient_ms_per_shot": round(timing["device_ms"] / timing["entries"], 4),
        "forward_ms": round(timing["forward_ms"], 2), "backward_ms": round(timing["backward_ms"], 2),
        "checkpoint_bytes": timing["snapshot_bytes"],
        "by_checkpoint_format": by_format,
        "roofline": {"bound": "hbm", "achieved": round(timing["bytes"] / (timing["device_ms"] * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(timing["bytes"] / (timing["device_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "k_batch_step",
                     "algorithmic_bytes_per_launch": round(timing["bytes"] / timing["launches"]),
                     "avg_launch_us": round(1e3 * timing["device_ms"] / timing["launches"], 2),
                     "launches": timing["launches"],
                     "tile_x": timing["tile_x"], "block_order": timing["block_order"],
                     "note": "BF16 checkpoints; forward and backward loops timed with HIP events on the batch's "
                             "stream; the active entries shrink as the shorter grids finish (112, 48, then 16); "
                             "tile_x, block_order: the step kernels' tile width and block order (1: a model's "
                             "shots tile by tile) picked by the batch's tuner"},
        "gradient_rel_l2_vs_fp64": res,
        "gradient_tolerance": 1e-3,
    }
    print(json.dumps(line))


def learned_workload(args):
    """Learned coarse-grid correction FWI against the plain fine FP32 FWI.

    The network is trained on the GPU (its own kernels) from fine 10th-order
    runs of 64 seeded random layered models on a 161x80 grid with the same
    spacing, time step and band as the test grid (it is a 7x7-cell local
    operator, so it applies to any grid of that spacing), its output gain
    picked on two validation models of that grid; or (`--train full`) from
    crops of fine runs of 2000 seeded models of the 501x1501 grid, its gain
    picked on two validation models of that grid.  Both report their
    seismogram error on the two 501x1501 validation models.  On a held-out 501x1501 model (12.5 m,
    nb=30, nt=3000, 8 Hz, `shots` surface shots, a receiver on every surface
    cell) `iters` FWI iterations run three ways: the plain fine FP32
    propagator, the 2x-coarser 4th-order propagator (25 m, 2 ms) with the
    network applied every 4 steps, and the same coarse propagator without
    it.  The fine propagator judges every final model (data misfit on the
    fine grid)."""
    from swfwi_amd import capi, learned, workloads
    from swfwi_amd.batch import Batch, Group

    learned.GRAPHS = not args.no_graphs
    t0 = time.perf_counter()
    case = learned.fwi_501x1501_case()
    val = learned.random_models(case, 2, 10 ** 6)  # validation models, seeds far from every training seed
    gains = (1.0, 0.5, 0.25)
    if args.train == "small":
        train_case = dict(nx=161, nz=80, nb=20, dx=12.5, dt=0.001, fm=8.0, nt=1200, vmin=1500.0, vmax=4500.0)
        res = learned.train_and_evaluate(train_case, n_train=64, seed=21)
        net = res["net"]
        training = {"models": 64, "grid": "161x80 at 12.5 m (whole fields)", "gain": res["gain"],
                    "held_out_seismogram_error_161x80": {"plain_coarse": round(res["err_plain"], 4),
                                                         "learned": round(res["err_learned"], 4)}}
        v_plain, (v_err,), _ = learned.held_out_errors(case, val, [net])
    else:  # 2000 models of the 501x1501 grid, crops; the output gain picked on the validation models
        raw, _ = learned.train_crops(case, 2000, seed=21, steps_per_round=6000, batch=32, floor=0.01,
                                     log=lambda m: print(m, file=sys.stderr, flush=True))
        cands = [learned.with_gain(raw, g) for g in gains]
        v_plain, errs, _ = learned.held_out_errors(case, val, cands)
        k = int(np.argmin(errs))
        net, v_err = cands[k], errs[k]
        training = {"models": 2000, "grid": "501x1501 at 12.5 m (112x112 coarse crops)", "gain": gains[k]}
    training["validation_seismogram_error_501x1501"] = {"plain_coarse": round(v_plain, 4), "learned": round(v_err, 4)}
    t_train = time.perf_counter() - t0
    training["seconds"] = round(t_train, 1)
    print(f"[bench] trained in {t_train:.1f} s: {training}", file=sys.stderr, flush=True)
    nx = case["nx"]
    shots = (np.linspace(0, nx - 1, args.shots).astype(np.int64) & ~1).astype(np.int32)
    fine, coarse = learned.grids(case, shots)
    vtrue = learned.random_models(case, 1, 9001)[0]  # [nx][nz], m/s
    v0 = workloads.smoothed(vtrue, 20.0)
    fb = Batch([Group(**fine)], snap_prec=capi.SNAP_NONE, order=capi.ORDER_10)
    fb.set_graphs(learned.GRAPHS)
    fb.bind_model(0, 0, learned.expand(vtrue, fine))
    dobs_f = np.stack(fb.split_data(fb.forward()))  # [shot][nt][ng]
    fb.close()
    dobs_c = learned.fine_on_coarse(dobs_f)
    iters = args.iters
    log = (lambda m: print(m, file=sys.stderr, flush=True))
    out = {}
    finals = {}
    runs = (("plain_fine_fp32", fine, capi.ORDER_10, v0, dobs_f, None),
            ("learned_coarse", coarse, capi.ORDER_4, learned.restrict(v0), dobs_c, net),
            ("plain_coarse", coarse, capi.ORDER_4, learned.restrict(v0), dobs_c, None))
    for name, grid, order, vstart, d, n in runs:
        t1 = time.perf_counter()
        print(f"[bench] {name}: {iters} FWI iterations", file=sys.stderr, flush=True)
        keep = tuple(k for k in (5, 10, 15) if k < iters) if grid is coarse else ()
        v, objs, gms, pts, kept = learned.fwi_run(grid, order, vstart, d, iters, case["vmin"], case["vmax"], net=n,
                                                  log=log, keep=keep)
        wall = time.perf_counter() - t1
        vf = v if grid is fine else learned.upsample(v, (nx, case["nz"]))
        misfit = learned.fine_misfit(case, fine, vf, dobs_f)
        out[name] = {"gradient_ms_per_shot": round(gms, 3), "gpts_per_s": round(pts / (gms * len(shots) * 1e-3) / 1e9, 2),
                     "fine_judged_misfit": float(f"{misfit:.6e}"), "own_grid_misfit": [float(f"{o:.5e}") for o in objs],
                     "wall_s": round(wall, 1)}
        finals[name] = (vf, objs, misfit, {k: learned.upsample(m, (nx, case["nz"])) for k, m in kept.items()})
    m0 = learned.fine_misfit(case, fine, v0, dobs_f)
    # equal misfit.  (1) The plain fine run's own objectives are the
    # fine-judged misfits of its successive models: the first of them at or
    # below the learned run's final misfit gives the fine iterations that
    # reach the same misfit.  (2) Fine fallback: from a coarse run's model
    # after n coarse iterations the fine propagator iterates on until the
    # misfit is within 2% of the plain fine run's final one; the schedule is
    # charged n coarse and k fine gradients.  Every schedule is reported, for
    # the learned run and for the coarse run without the network (the control).
    _, objs_f, m_plain, _ = finals["plain_fine_fp32"]
    m_learned = finals["learned_coarse"][2]
    seq = list(objs_f) + [m_plain]  # misfit after k = 0 .. iters fine updates
    k_eq = next((k for k, m in enumerate(seq) if m <= m_learned), None)
    t_f, t_l = out["plain_fine_fp32"]["gradient_ms_per_shot"], out["learned_coarse"]["gradient_ms_per_shot"]
    target = 1.02 * m_plain
    schedules = {}
    for name in ("learned_coarse", "plain_coarse"):
        t_m = out[name]["gradient_ms_per_shot"]
        starts = dict(finals[name][3])
        starts[iters] = finals[name][0]
        rows = []
        for n_c in sorted(starts):
            print(f"[bench] fine fallback from {name} after {n_c} iterations", file=sys.stderr, flush=True)
            _, objs_h, _, _, _ = learned.fwi_run(fine, capi.ORDER_10, starts[n_c], dobs_f, iters, case["vmin"],
                                                 case["vmax"], stop_at=target)
            k = len(objs_h) - 1 if objs_h and objs_h[-1] <= target else None
            rows.append({"coarse_iterations": n_c, "fine_iterations": k,
                         "speedup": round(iters * t_f / (n_c * t_m + k * t_f), 3) if k is not None else None})
        schedules[name] = rows
    best = {n: max((r for r in rows if r["speedup"]), key=lambda r: r["speedup"], default=None)
            for n, rows in schedules.items()}
    equal = {
        "plain_fine_iterations_to_learned_misfit": k_eq,
        "speedup_to_learned_misfit": round(k_eq * t_f / (iters * t_l), 3) if k_eq else None,
        "fallback_schedules": schedules,
        "best_learned_with_fallback": best["learned_coarse"],
        "best_control_with_fallback": best["plain_coarse"],
        "note": "speedup = plain fine FP32 gradient time per shot over its iterations / the schedule's (coarse "
                "iterations x coarse gradient time + fine iterations x fine gradient time), at a misfit within 2% "
                "of the plain fine run's final one"}
    # the network kernel alone: one launch over the 16 coarse fields of a
    # correction (the forward of the learned run holds 2 per 4 steps)
    from scipy.ndimage import gaussian_filter
    cshape = (coarse["nx"] + 2 * coarse["nb"] + 12, coarse["nz"] + coarse["nb"] + 12)
    rng = np.random.default_rng(1)
    fields = np.stack([gaussian_filter(rng.standard_normal(cshape), 3.0) for _ in shots]).astype(np.float32)
    rvc = np.broadcast_to(1.0 / learned.expand(learned.restrict(v0), coarse), fields.shape).astype(np.float32)
    reg = learned.region(coarse)
    net_ms = net.time_apply(fields, rvc, reg, reps=50)
    npts = (reg[1] - reg[0]) * (reg[3] - reg[2]) * len(shots)
    flops = 2.0 * (9 * 2 * 16 + 9 * 16 * 16 + 9 * 16) * npts  # multiply-adds of the three layers per point
    tfs = flops / (net_ms * 1e-3) / 1e12
    pf, lc, pc = out["plain_fine_fp32"], out["learned_coarse"], out["plain_coarse"]
    speedup = pf["gradient_ms_per_shot"] / lc["gradient_ms_per_shot"]
    ratio = lc["fine_judged_misfit"] / pf["fine_judged_misfit"]
    line = {
        "metric": "FWI gradient time/shot & Gpts/s, learned vs plain FP32 at equal data misfit",
        "value": round(speedup, 3), "unit": "x (plain fine FP32 gradient time/shot / learned coarse)",
        "n_gpus": 1, "steps": iters, "warmup": 0, "ms_per_step": round(1e3 * lc["wall_s"] / iters, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (network: bf16 MFMA)",
        "data": "synthetic (seeded random layered models; observed data from the fine propagator on the true "
                "held-out model)",
        "config": {"workload": f"FWI {iters} iterations on a held-out 501x1501 model (12.5 m, nb=30, nt=3000, 8 Hz, "
                               f"{len(shots)} shots x 1501 receivers): plain fine 10th-order FP32 vs 2x-coarser "
                               "4th-order (25 m, 2 ms) with the 3-layer 3x3 16-channel correction every 4 steps",
                   "parallelism": "single GPU"},
        "learned_vs_plain": {
            "gradient_speedup": round(speedup, 3),
            "fine_misfit_ratio_learned_over_plain": round(ratio, 4),
            "equal_misfit_within_2pct": bool(abs(ratio - 1.0) <= 0.02),
            "fine_misfit_initial_model": float(f"{m0:.6e}"),
            "plain_coarse_fine_misfit_ratio": round(pc["fine_judged_misfit"] / pf["fine_judged_misfit"], 4),
            "at_equal_misfit": equal,
        },
        "roofline": {"bound": "mfma", "achieved": round(tfs, 1), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(tfs / 2500.0, 4), "traffic": None, "kernel": "k_net_apply",
                     "algorithmic_flop_per_launch": flops, "avg_launch_us": round(1e3 * net_ms, 2),
                     "note": f"{len(shots)} coarse fields of {cshape[0]}x{cshape[1]} (region {npts // len(shots)} "
                             "points each), 5472 algorithmic FLOP per point (2736 multiply-adds), BF16 MFMA with "
                             "FP32 accumulation; 50 back-to-back launches timed with HIP events"},
        "runs": out,
        "training": training,
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["exact", "fp32"], default="exact")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-epoch", action="store_true", help="also time a whole plain-FWI epoch on the CPU")
    ap.add_argument("--nt", type=int, default=None, help="override nt (testing only)")
    ap.add_argument("--workload", choices=["ess_301x921", "fwi_201x301", "mixed", "learned"], default="ess_301x921")
    ap.add_argument("--shots", type=int, default=16, help="shots per model (mixed and learned workloads)")
    ap.add_argument("--iters", type=int, default=20, help="FWI iterations (learned workload)")
    ap.add_argument("--train", choices=["small", "full"], default="small",
                    help="learned workload: train on 64 whole 161x80 grids (default) or on crops of 2000 501x1501 "
                         "fine runs")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the time loops directly instead of replaying HIP graphs (for kernel tracers)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (shots/members shard across GPUs "
              "with one gradient all-reduce per iteration; that path is not implemented)", file=sys.stderr)
        sys.exit(2)

    from swfwi_amd import capi, workloads

    if capi.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    if args.workload == "mixed":
        return mixed_workload(args)
    if args.workload == "learned":
        return learned_workload(args)
    ess = args.workload == "ess_301x921"
    case = workloads.ess_301x921(nt=args.nt or 4000) if ess else workloads.fwi_201x301(nt=args.nt or 2000)
    prec = capi.PREC_EXACT if args.precision == "exact" else capi.PREC_FP32
    vtrue, v0 = workloads.models(case)
    vt = capi.expand_domain(vtrue, case.nb, case.dx, case.dt)
    vi = capi.expand_domain(v0, case.nb, case.dx, case.dt)
    wlt = capi.ricker(case.nt, case.fm, case.dt, case.amp)
    s = capi.Solver(case.nx, case.nz, case.nb, case.dt, case.dx, case.fm, case.nt, case.sx, case.sz, case.gx,
                    case.gz, prec)
    # observed data: every shot on the true model, batched (fm-damp)
    s.bind_velocity(vt)
    # (its time loop is timed like the epochs' loops: the batched forward step
    # over all shots is the HBM-streaming kernel of this workload)
    s.set_timing(True)
    pu_obs0 = s.counters()[0][0]
    t0 = time.perf_counter()
    dobs = s.forward_shots(wlt)
    t_obs = time.perf_counter() - t0
    obs_ms, obs_n = s.kernel_times()["fwd"]
    obs_pu = s.counters()[0][0] - pu_obs0
    s.set_timing(False)
    print(f"[bench] observed data {case.ns} shots in {t_obs:.2f} s", file=sys.stderr, flush=True)
    s.bind_velocity(vi)
    inv = capi.Inversion(s, wlt, dobs, case.vmin, case.vmax, 5, 200.0)
    if args.no_graphs:
        s.set_graphs(False)
    objs = []
    epoch = inv.ess_epoch if ess else inv.fwi_epoch
    for it in range(args.warmup):
        epoch(it)  # also captures the time loops into HIP graphs
        objs.append(inv.stats()["obj1"])
        print(f"[bench] warmup epoch {it} obj {objs[-1]:.4e}", file=sys.stderr, flush=True)
    # the timed epochs, with one HIP event pair around every time loop on the
    # library's stream (outside the replayed graphs): the per-launch average
    # of the dominant step kernel is its loop time over its launches
    s.set_timing(True)
    pu0 = s.counters()[0]
    stats = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        epoch(args.warmup + k)
        stats.append(inv.stats())
    t1 = time.perf_counter()
    pu1 = s.counters()[0]
    kt = s.kernel_times()
    s.set_timing(False)
    for st in stats:
        objs.append(st["obj1"])
    wall = t1 - t0
    gpts = (sum(pu1) - sum(pu0)) / wall / 1e9
    # dominant kernel: the one with the larger total time in the timed region
    fwd_ms, fwd_n = kt["fwd"]
    bwd_ms, bwd_n = kt["bwd"]
    grad_ms = float(np.mean([st["grad_ms"] for st in stats]))
    ls_ms = float(np.mean([st["linesearch_ms"] for st in stats]))
    ntrial = float(np.mean([st["ntrial"] for st in stats]))
    # algorithmic bytes from the E-weighted point updates of each kernel type
    # (the backward kernel advances two fields per point: 40 B per point and
    # backward iteration)
    fwd_pu = pu1[0] - pu0[0]
    bwd_pu = (pu1[1] - pu0[1]) / 2.0
    if bwd_ms >= fwd_ms:
        name, ms, n, byts = "k_bwd_step", bwd_ms, bwd_n, bwd_pu * BWD_BYTES_PER_PT
    else:
        name, ms, n, byts = "k_fwd_step", fwd_ms, fwd_n, fwd_pu * FWD_BYTES_PER_PT
    achieved = byts / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(name, args.workload), "kernel": name,
            "algorithmic_bytes_per_launch": round(byts / max(n, 1)),
            "avg_launch_us": round(1e3 * ms / max(n, 1), 3), "launches": n,
            "tile_x": s.tile_x(0 if name == "k_fwd_step" else 1, 1) if ess else None,
            "note": ("one-entry launches (the two line-search trials as two loops side by side on two streams, "
                     "each launch timed on its own stream)" if ess else "batches of 11-22 small-grid entries") +
                    ": the live set fits the 256 MB last-level cache, so achieved is an effective (algorithmic) "
                    "bandwidth of latency-bound launches; traffic is the measured FETCH_SIZE + WRITE_SIZE per "
                    "launch"}

    obs_bytes = obs_pu * FWD_BYTES_PER_PT
    obs_achieved = obs_bytes / (obs_ms * 1e-3) / 1e9 if obs_ms > 0 else 0.0
    roof_batched = {"bound": "hbm", "achieved": round(obs_achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(obs_achieved / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic("k_fwd_step_batched", args.workload), "kernel": "k_fwd_step",
                    "entries": case.ns, "algorithmic_bytes_per_launch": round(obs_bytes / max(obs_n, 1)),
                    "avg_launch_us": round(1e3 * obs_ms / max(obs_n, 1), 3), "launches": obs_n,
                    "tile_x": s.tile_x(0, case.ns),
                    "note": "the observed-data generation before the timed epochs (fm-damp): every shot in one "
                            "batched forward loop, untimed for value; algorithmic bytes charge the shared model "
                            "per shot"}

    cpu = None
    if not args.no_cpu_baseline:
        res, threads = cpu_baseline(case, ess)
        ref = res.get("reference_threading")
        par = res.get("all_parallel")
        what = "ESS calgradient (one supershot) on the 301x921 grid" if ess else \
            "FWI calgradient (one shot) on the 201x301 grid"
        cpu = {"value": round(ref[0], 4), "unit": "Gpts/s", "cores": threads, "kind": "port",
               "sample": f"oracle {what}, nt={ref[2]} ({ref[2]} forward steps + {ref[3]} backward iterations), "
                         f"{ref[1]:.1f} s, serial u2 pass as the reference build",
               "all_parallel_value": round(par[0], 4) if par else None}
        if ess or args.cpu_epoch:
            # the whole epoch: the ESS one (~18 s) by default, the plain-FWI
            # one (11 shot gradients + 22 trial runs, ~25 s) on request
            t_ep, ntr = cpu_epoch(case, ess, dobs)
            grad_s = ref[1] * (1 if ess else case.ns)
            cpu["epoch"] = {"seconds": round(t_ep, 2), "gradient_seconds": round(grad_s, 2),
                            "linesearch_seconds": round(max(t_ep - grad_s, 0.0), 2), "trial_forward_runs": ntr,
                            "gpu_epoch_seconds": round(wall / args.steps, 4),
                            "note": "one oracle epoch with the reference's threading; gradient_seconds from the "
                                    "gradient timed above" + ("" if ess else " times the shots") +
                                    ", line search = the rest (trial forward runs, objectives, update)"}

    line = {
        "metric": "FWI gradient time/shot & Gpts/s, learned vs plain FP32 at equal data misfit",
        "value": round(gpts, 3),
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded layered model; observed data from the same propagator on the true model)",
        "config": {"workload": (f"ESS-FWI 301x921 dx=12.5 nb=30 nt={case.nt}, 60 shots x 921 receivers, one "
                                "supershot per epoch") if ess else
                               (f"FWI 201x301 dx=10 nb=20 nt={case.nt}, 11 shots x 301 receivers, every shot's "
                                "gradient per epoch"),
                   "precision_mode": args.precision, "graphs": not args.no_graphs, "parallelism": "single GPU"},
        ("gradient_ms_per_supershot" if ess else "gradient_ms_per_shot"): round(grad_ms if ess else
                                                                             grad_ms / case.ns, 3),
        "linesearch_ms": round(ls_ms, 3),
        "trial_forward_runs_per_epoch": ntrial,
        "misfit": [float(f"{o:.6e}") for o in objs],
        "roofline": roof,
        "roofline_batched": roof_batched,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
