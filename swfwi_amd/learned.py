"""Learned coarse-grid correction: data generation from the fine propagator,
training on the GPU, and the coarse propagator with the correction.

The fine grid (10th-order stencil, dx, dt) and the coarse grid (4th-order
stencil, 2 dx, 2 dt, half the sponge) share the physical model: coarse cell
(i, k) is fine cell (2i, 2k).  Training pairs come from fine runs of random
layered models: the coarse leapfrog is started from the fine wavefield at
coarse steps j-1, j (fine steps 2j-2, 2j), run 4 coarse steps, and the
network learns to map the result at coarse steps j+3, j+4 onto the fine
wavefield there (fine steps 2j+6, 2j+8).  The coarse forward runs apply the
correction every 4 steps; the fine propagator judges the result."""
import numpy as np

from . import capi, workloads
from .batch import Batch, Group, Net


def small_case():
    """A test-sized fine grid: 161 x 80 cells of 10 m, nb=20, dt=1 ms,
    nt=1200, 8 Hz Ricker at the centre (depth 3 cells), receivers on row 1;
    velocities 1500-3500 m/s."""
    return dict(nx=161, nz=80, nb=20, dx=10.0, dt=0.001, fm=8.0, nt=1200, vmin=1500.0, vmax=3500.0)


def fwi_501x1501_case():
    """The 501 x 1501 fine grid at 12.5 m (nb=30, dt=1 ms, 8 Hz, nt=3000)
    and its 250 x 751 coarse grid at 25 m (nb=15, dt=2 ms)."""
    return dict(nx=1501, nz=501, nb=30, dx=12.5, dt=0.001, fm=8.0, nt=3000, vmin=1500.0, vmax=4500.0)


# Coarse cell (i, k) is fine cell (2i, 2k + 1): the zero rows above the
# first row of either grid (the free surface of the reference's scheme) then
# lie at the same depth, two fine cells above fine row 1.
OZ = 1


# Whether the batches this module creates replay their time loops as HIP
# graphs (the default) or launch every step directly (for kernel tracers).
GRAPHS = True


def _batch(grid, prec, order):
    b = Batch([Group(**grid)], snap_prec=prec, order=order)
    b.set_graphs(GRAPHS)
    return b


def grids(case, shots=None):
    """(fine, coarse) Group keyword dicts of one shot per model.  shots: fine
    physical source x positions (even; default the centre); the sources sit
    on fine row 3 (coarse row 1), the receivers on fine row 1 (coarse row 0)
    at every coarse column."""
    nx, nz, nb = case["nx"], case["nz"], case["nb"]
    if nb % 2 or nx % 2 == 0:
        raise ValueError("the fine grid needs an odd nx and an even nb")
    sx = np.array(shots if shots is not None else [(nx // 2) & ~1], np.int32)
    if np.any(sx % 2):
        raise ValueError("fine source positions must be even")
    cx, cz = (nx + 1) // 2, (nz + 1 - OZ) // 2
    fine = dict(nx=nx, nz=nz, nb=nb, dx=case["dx"], dt=case["dt"], fm=case["fm"], nt=case["nt"], sx=sx,
                sz=np.full(sx.size, 2 + OZ, np.int32), gx0=0, ng=nx, gz=OZ, nmodels=1)
    # the coarse sponge: half the cells, twice the step, so twice the per-step
    # damping for the same attenuation per unit time
    coarse = dict(nx=cx, nz=cz, nb=nb // 2, dx=2 * case["dx"], dt=2 * case["dt"], fm=case["fm"],
                  nt=case["nt"] // 2, sx=sx // 2, sz=np.full(sx.size, 1, np.int32), gx0=0, ng=cx, gz=0, nmodels=1,
                  dscale=2.0)
    return fine, coarse


def random_models(case, n, seed):
    """n seeded random layered models on the fine grid (the layered recipe of
    workloads.layered_model),
    [n][nx][nz] in m/s."""
    return np.stack([workloads.layered_model(case["nz"], case["nx"], case["vmin"], case["vmax"], seed + k)
                     for k in range(n)])


def restrict(v):
    """Physical fine model -> coarse model: fine cell (2i, 2k + 1)."""
    v = np.asarray(v)
    cz = (v.shape[1] + 1 - OZ) // 2
    return np.ascontiguousarray(v[::2, OZ::2][:, :cz])


def expand(v, grid):
    return capi.expand_domain(np.ascontiguousarray(v, np.float32), grid["nb"], grid["dx"], grid["dt"])


def region(grid):
    """The physical interior of a padded grid: (x0, x1, z0, z1)."""
    nxpad, nzpad, bx0 = grid["nx"] + 2 * grid["nb"] + 12, grid["nz"] + grid["nb"] + 12, grid["nb"] + 6
    return (bx0, nxpad - bx0, 6, nzpad - bx0)


def with_models(grid, n):
    g = dict(grid)
    g["nmodels"] = n
    return g


def training_pairs(case, models, times):
    """Inputs (the coarse leapfrog after 4 steps from the fine wavefield) and
    targets (the fine wavefield there) for every model and coarse start step j
    in `times`: [2 * len(models) * len(times)][nxpad'][nzpad'] each, and the
    fine seismograms [model][nt][ng]."""
    fine, coarse = grids(case)
    n = len(models)
    fb = _batch(with_models(fine, n), capi.SNAP_NONE, capi.ORDER_10)
    for m in range(n):
        fb.bind_model(0, m, expand(models[m], fine))
    taps = []
    for j in times:
        taps += [2 * j - 2, 2 * j, 2 * j + 6, 2 * j + 8]
    fb.set_taps(taps, oz=OZ)
    dfine = np.stack(fb.split_data(fb.forward()))  # [model][nt][ng]
    T = np.stack(fb.taps())  # [model][tap][x][z]
    fb.close()
    cb = _batch(with_models(coarse, n), capi.SNAP_NONE, capi.ORDER_4)
    for m in range(n):
        cb.bind_model(0, m, expand(restrict(models[m]), coarse))
    ins, tgts = [], []
    shape = T.shape[2:]
    for k, j in enumerate(times):
        prev, cur = cb.step_from(T[:, 4 * k], T[:, 4 * k + 1], it0=j, nsteps=4)
        ins += [prev.reshape((n,) + shape), cur.reshape((n,) + shape)]
        tgts += [T[:, 4 * k + 2], T[:, 4 * k + 3]]
    cb.close()
    rv = np.stack([1.0 / expand(restrict(models[m]), coarse) for m in range(n)]).astype(np.float32)
    rvs = np.concatenate([rv] * (2 * len(times)))
    return (np.ascontiguousarray(np.concatenate(ins), np.float32), np.ascontiguousarray(rvs),
            np.ascontiguousarray(np.concatenate(tgts), np.float32), dfine)


def coarse_seismograms(case, models, net=None):
    """Coarse forward runs of the models (with the correction when net is
    given): [model][nt'][ng']."""
    _, coarse = grids(case)
    n = len(models)
    cb = _batch(with_models(coarse, n), capi.SNAP_NONE, capi.ORDER_4)
    for m in range(n):
        cb.bind_model(0, m, expand(restrict(models[m]), coarse))
    if net is not None:
        cb.set_correction(net)
    d = cb.forward()
    out = np.stack(cb.split_data(d))
    cb.close()
    return out


def fine_on_coarse(dfine):
    """Fine seismograms [model][nt][ng] at the coarse samples and receivers."""
    return np.ascontiguousarray(np.asarray(dfine)[:, ::2, ::2])


def seis_error(dc, df):
    return float(np.linalg.norm(dc - df) / np.linalg.norm(df))


def rollout_pairs(case, models, starts, rounds, net=None):
    """Training pairs seen along the corrected propagation.  For every model
    and start step j (j % 4 == 3, so that a correction falls after every
    4th step as in the forward runs) and round r < rounds: the coarse run
    from the fine wavefield at (j-1, j) is corrected by `net` for 4r steps
    (none for r = 0), then 4 plain steps give the input; the fine wavefield
    at (j+4r+3, j+4r+4) is the target.  Returns inputs, their 1 / vel_t,
    targets ([2 * models * starts * rounds][nxpad'][nzpad']) and the fine
    seismograms."""
    fine, coarse = grids(case)
    n = len(models)
    if any(j % 4 != 3 for j in starts):
        raise ValueError("rollout starts must be 3 mod 4")
    fb = _batch(with_models(fine, n), capi.SNAP_NONE, capi.ORDER_10)
    for m in range(n):
        fb.bind_model(0, m, expand(models[m], fine))
    taps, slot = [], {}
    for j in starts:
        for k in [j - 1, j] + [j + 4 * r + d for r in range(rounds) for d in (3, 4)]:
            if k not in slot:
                slot[k] = len(taps)
                taps.append(2 * k)
    fb.set_taps(taps, oz=OZ)
    dfine = np.stack(fb.split_data(fb.forward()))
    T = np.stack(fb.taps())  # [model][tap][x][z]
    fb.close()
    cb = _batch(with_models(coarse, n), capi.SNAP_NONE, capi.ORDER_4)
    for m in range(n):
        cb.bind_model(0, m, expand(restrict(models[m]), coarse))
    shape = T.shape[2:]
    ins, tgts = [], []
    for r in range(rounds):
        for j in starts:
            prev, cur = T[:, slot[j - 1]], T[:, slot[j]]
            if r > 0:
                cb.set_correction(net)
                prev, cur = cb.step_from(prev, cur, it0=j, nsteps=4 * r)
                cb.set_correction(None)
            prev, cur = cb.step_from(prev, cur, it0=j + 4 * r, nsteps=4)
            ins += [prev.reshape((n,) + shape), cur.reshape((n,) + shape)]
            tgts += [T[:, slot[j + 4 * r + 3]], T[:, slot[j + 4 * r + 4]]]
    cb.close()
    rv = np.stack([1.0 / expand(restrict(models[m]), coarse) for m in range(n)]).astype(np.float32)
    rvs = np.concatenate([rv] * (len(ins)))[:len(ins) * n]
    return (np.ascontiguousarray(np.concatenate(ins), np.float32), np.ascontiguousarray(rvs, np.float32),
            np.ascontiguousarray(np.concatenate(tgts), np.float32), dfine)


def _noisy(ins, sigma, copies, rng):
    """`copies` copies of the inputs, each with white noise of `sigma` times
    the input's own RMS added (the targets stay clean, so the network also
    learns to take such perturbations out again: the corrected propagation
    stays stable where it meets states the clean pairs never showed)."""
    rms = np.sqrt(np.mean(np.square(ins, dtype=np.float64), axis=(1, 2)))[:, None, None]
    return np.ascontiguousarray(np.concatenate(
        [ins + (sigma * rms * rng.standard_normal(ins.shape)).astype(np.float32) for _ in range(copies)]), np.float32)


def train(case, models, rounds=3, steps_per_round=3000, starts=None, batch=16, lr=1e-3, noise=0.02, copies=2,
          seed=0, net=None, log=None):
    """Train the correction on the given fine models.  Round 0 uses pairs
    from the fine wavefield; each later round r adds the pairs met along the
    current network's own corrected propagation (4 r corrected steps, then
    4 plain ones) and trains on everything gathered so far.  Every input is
    used in `copies` noisy copies (`noise` times its RMS).  Returns (net,
    losses per round)."""
    _, coarse = grids(case)
    nt_c = case["nt"] // 2
    if starts is None:
        starts = list(range(11, nt_c - 4 * rounds - 8, 8))  # all 3 mod 4
    reg = region(coarse)
    rng = np.random.default_rng(seed + 7)
    net = net or Net(Net.initial(seed=seed, scale=1e-3))
    data = None
    hist = []
    for r in range(rounds):
        ins, rvs, tgts, _ = rollout_pairs(case, models, starts, r + 1, net if r else None)
        # keep only the newest round's pairs from this call (earlier ones repeat)
        per = 2 * len(models) * len(starts)
        new = (_noisy(ins[-per:], noise, copies, rng), np.concatenate([rvs[-per:]] * copies),
               np.concatenate([tgts[-per:]] * copies))
        data = new if data is None else tuple(np.concatenate([a, b]) for a, b in zip(data, new))
        data = tuple(np.ascontiguousarray(a, np.float32) for a in data)
        net.set_data(*data, reg)
        losses = net.train(batch=min(batch, len(data[0])), steps=steps_per_round, lr=lr, seed=seed + r)
        hist.append(losses)
        if log:
            log(f"round {r}: {len(data[0])} pairs, loss {losses[0]:.3e} -> {losses[-50:].mean():.3e}")
    return net, hist


def crop_pairs(case, n_models, rnd, net=None, seed=0, chunk=100, starts=2, crop=112, crops=2, floor=0.1, log=None):
    """Training pairs of round `rnd` from full-size fine runs, as crops.

    n_models seeded random layered models (`seed + k`, workloads.layered_model)
    in chunks of `chunk`: each chunk is one fine batch (one shot per model, at
    a seeded even x on fine row 3, the same for the chunk) tapped at the
    coarse points; from each of `starts` seeded coarse start steps j (3 mod 4)
    the coarse leapfrog runs 4 rnd corrected steps and then 4 plain ones from
    the fine wavefield at (j-1, j), and the fine wavefield at (j+4rnd+3,
    j+4rnd+4) is the target, as in rollout_pairs.  From every (model, start,
    time level) `crops` windows of crop x crop coarse cells are cut, centred
    on cells where the target's magnitude is at least `floor` times its maximum
    (the network is a 7x7-cell local operator: a window with a 5-cell margin
    trains it as the whole field would).  Returns inputs, 1 / vel_t and
    targets [n][crop][crop]."""
    nt_c = case["nt"] // 2
    ins, rvs, tgts = [], [], []
    for c0 in range(0, n_models, chunk):
        m = min(chunk, n_models - c0)
        rng = np.random.default_rng(seed + 7919 * (c0 // chunk + 1))
        nx = case["nx"]
        sx = int(rng.integers(nx // 8, 7 * nx // 8)) & ~1
        js = sorted(int(j) for j in rng.choice(np.arange(11, nt_c - 4 * rnd - 8, 4), size=starts, replace=False))
        models = random_models(case, m, seed + c0)
        fine, coarse = grids(case, [sx])
        slot, taps = {}, []
        for j in js:
            for k in (j - 1, j, j + 4 * rnd + 3, j + 4 * rnd + 4):
                if k not in slot:
                    slot[k] = len(taps)
                    taps.append(2 * k)
        fb = _batch(with_models(fine, m), capi.SNAP_NONE, capi.ORDER_10)
        for q in range(m):
            fb.bind_model(0, q, expand(models[q], fine))
        fb.set_taps(taps, oz=OZ)
        fb.forward(want_data=False)
        T = np.stack(fb.taps())  # [model][tap][x][z]
        fb.close()
        cb = _batch(with_models(coarse, m), capi.SNAP_NONE, capi.ORDER_4)
        rv = []
        for q in range(m):
            vc = expand(restrict(models[q]), coarse)
            cb.bind_model(0, q, vc)
            rv.append((1.0 / vc).astype(np.float32))
        shape = T.shape[2:]
        for j in js:
            prev, cur = T[:, slot[j - 1]], T[:, slot[j]]
            if rnd > 0:
                cb.set_correction(net)
                prev, cur = cb.step_from(prev, cur, it0=j, nsteps=4 * rnd)
                cb.set_correction(None)
            prev, cur = cb.step_from(prev, cur, it0=j + 4 * rnd, nsteps=4)
            pairs = ((prev.reshape((m,) + shape), T[:, slot[j + 4 * rnd + 3]]),
                     (cur.reshape((m,) + shape), T[:, slot[j + 4 * rnd + 4]]))
            for q in range(m):
                for inp, tgt in pairs:
                    t = tgt[q]
                    mag = np.abs(t)
                    cand = np.argwhere(mag >= floor * mag.max()) if mag.max() > 0 else np.argwhere(mag >= 0)
                    for ci in rng.choice(len(cand), size=crops):
                        x0 = int(np.clip(cand[ci][0] - crop // 2, 0, shape[0] - crop))
                        z0 = int(np.clip(cand[ci][1] - crop // 2, 0, shape[1] - crop))
                        w = (slice(x0, x0 + crop), slice(z0, z0 + crop))
                        ins.append(inp[q][w])
                        rvs.append(rv[q][w])
                        tgts.append(t[w])
        cb.close()
        if log:
            log(f"  round {rnd}: models {c0 + m}/{n_models}, {len(ins)} crops")
    f = lambda a: np.ascontiguousarray(np.stack(a), np.float32)  # noqa: E731
    return f(ins), f(rvs), f(tgts)


def train_crops(case, n_models, rounds=3, steps_per_round=3000, batch=16, lr=1e-3, noise=0.02, copies=2, seed=0,
                log=None, **kw):
    """train() on crop_pairs: round 0 from the fine wavefield, later rounds
    from the network's own corrected propagation, noise-augmented inputs,
    everything gathered so far trained on.  Returns (net, losses per round)."""
    rng = np.random.default_rng(seed + 7)
    net = Net(Net.initial(seed=seed, scale=1e-3))
    data, hist = None, []
    for r in range(rounds):
        ins, rvs, tgts = crop_pairs(case, n_models, r, net if r else None, seed=seed, log=log, **kw)
        new = (_noisy(ins, noise, copies, rng), np.concatenate([rvs] * copies), np.concatenate([tgts] * copies))
        data = new if data is None else tuple(np.concatenate([a, b]) for a, b in zip(data, new))
        data = tuple(np.ascontiguousarray(a, np.float32) for a in data)
        c = ins.shape[1]
        net.set_data(*data, (5, c - 5, 5, c - 5))
        losses = net.train(batch=min(batch, len(data[0])), steps=steps_per_round, lr=lr, seed=seed + r)
        hist.append(losses)
        if log:
            log(f"round {r}: {len(data[0])} crops, loss {losses[0]:.3e} -> {losses[-50:].mean():.3e}")
    return net, hist


def with_gain(net, gain):
    """The network with its output scaled by `gain` (the last layer's
    weights times gain; the network is linear in them)."""
    w = net.weights()
    w[-9 * 16:] *= gain
    return Net(w)


def held_out_errors(case, models, nets):
    """Seismogram relative L2 errors against the fine propagation, at the
    coarse samples and receivers, of the plain coarse propagation and of the
    coarse propagation with each network of `nets`."""
    _, _, _, dfh = rollout_pairs(case, models, [11], 1)
    ref = fine_on_coarse(dfh)
    plain = coarse_seismograms(case, models)
    return seis_error(plain, ref), [seis_error(coarse_seismograms(case, models, n), ref) for n in nets], ref


def train_and_evaluate(case, n_train, n_val=2, n_test=2, gains=(1.0, 0.5, 0.25), seed=0, verbose=False, **kw):
    """Train a network on n_train random models, pick its output gain on
    n_val validation models (the one whose corrected coarse seismograms are
    closest to the fine ones) and compare the coarse propagation with and
    without it against the fine propagation on n_test further held-out
    models (seismogram relative L2 error at the coarse samples)."""
    models = random_models(case, n_train + n_val + n_test, seed)
    net, hist = train(case, models[:n_train], seed=seed, log=(lambda m: print(m, flush=True)) if verbose else None,
                      **kw)
    cands = [with_gain(net, g) for g in gains]
    val_plain, val_errs, _ = held_out_errors(case, models[n_train:n_train + n_val], cands)
    k = int(np.argmin(val_errs))
    test = models[n_train + n_val:]
    err_plain, (err_learned,), ref = held_out_errors(case, test, [cands[k]])
    for i, c in enumerate(cands):
        if i != k:
            c.close()
    if verbose:
        print(f"validation: plain {val_plain:.4f}, gains {dict(zip(gains, np.round(val_errs, 4)))}; "
              f"test: plain {err_plain:.4f} learned {err_learned:.4f}", flush=True)
    return {"err_plain": err_plain, "err_learned": err_learned, "gain": gains[k], "val_plain": val_plain,
            "val_errs": val_errs, "losses": hist, "net": cands[k]}


# ---- FWI with the plain fine propagator and with the learned coarse one ----

def _mask_scale(g, grid):
    """scaleGradient + maskGradient (damp4t10d.cpp:309-319,216-226) on a
    padded gradient: zero outside the physical interior, times sqrt(iz) of
    the physical depth inside (1 at the surface row)."""
    x0, x1, z0, z1 = region(grid)
    out = np.zeros_like(g)
    iz = np.arange(z1 - z0, dtype=np.float32)
    iz[0] = 1.0
    out[x0:x1, z0:z1] = g[x0:x1, z0:z1] * np.sqrt(iz)[None, :]
    return out


def _refill(vt, grid):
    """refillBoundary (damp4t10d.cpp:228-252): edge-replicate the interior."""
    x0, x1, z0, z1 = region(grid)
    ix = np.clip(np.arange(vt.shape[0]), x0, x1 - 1)
    iz = np.clip(np.arange(vt.shape[1]), z0, z1 - 1)
    return np.ascontiguousarray(vt[np.ix_(ix, iz)])


def fwi_run(grid, order, v0, dobs, iters, vmin, vmax, net=None, dv=30.0, snap_prec=None, log=None, keep=(),
            stop_at=None):
    """`iters` FWI iterations on one grid: every shot's gradient in one
    batched forward (checkpoints) + adjoint source + backward, PR+ conjugate
    gradient, a two-trial parabolic step on the summed objective, clamped
    update of the transformed model and boundary refill.  dobs: [shot][nt][ng]
    on this grid's receivers and samples.  Returns the final physical model,
    the objectives, the gradient time per shot (forward + backward on the
    device, ms; median over the iterations), the stencil point updates of
    one gradient of all shots, and {k: physical model after k updates} for
    the k in `keep`.  With `stop_at`, the run ends at the first model whose
    objective is at or below it (the objectives then end with that one)."""
    prec = capi.SNAP_BF16 if snap_prec is None else snap_prec
    S = len(grid["sx"])
    b = _batch(grid, prec, order)
    if net is not None:
        b.set_correction(net)
    tr = _batch(grid, capi.SNAP_NONE, order)  # trial forward runs
    if net is not None:
        tr.set_correction(net)
    dx, dt = grid["dx"], grid["dt"]
    tmin, tmax = (dx / dt / vmax) ** 2, (dx / dt / vmin) ** 2
    vt = expand(v0, grid)
    obs = np.ascontiguousarray(np.asarray(dobs, np.float32).ravel())

    def objective(model, batch):
        batch.bind_model(0, 0, model)
        d = batch.forward()
        return float(np.sum((obs.astype(np.float64) - d) ** 2)), d

    objs, gms, pts = [], [], 0.0
    g_prev = d_prev = None
    x0, x1, z0, z1 = region(grid)

    def physical(m):
        v = dx / (dt * np.sqrt(m.astype(np.float64)))
        return np.ascontiguousarray(v[x0:x1, z0:z1], np.float32)

    kept = {}
    for it in range(iters):
        b.bind_model(0, 0, vt)
        dcal = b.forward()
        J = float(np.sum((obs.astype(np.float64) - dcal) ** 2))
        if stop_at is not None and J <= stop_at:
            objs.append(J)
            break
        b.adjoint_source(obs, want=False)
        _, gm = b.backward(None, want_entries=False)
        info = b.info()
        gms.append((info["forward_ms"] + info["backward_ms"]) / S)
        pts = info["forward_points"] + 2 * info["backward_points"]  # the backward advances two fields
        g = _mask_scale(gm.reshape(vt.shape), grid)
        if d_prev is None:
            d = g
        else:  # PR+ (fwiframework.cpp:42-72)
            beta = max(0.0, float(np.sum(g * (g - g_prev)) / max(np.sum(g_prev * g_prev), 1e-30)))
            d = g + beta * d_prev
        g_prev, d_prev = g, d
        # step: the largest velocity change of alpha * d is dv m/s
        v = dx / (dt * np.sqrt(vt))
        dvdt = v / (2 * vt)  # |d v / d vel_t|
        a0 = dv / max(float(np.max(np.abs(d) * dvdt)), 1e-30)
        trial = []
        for a in (a0, 2 * a0):
            m = _refill(np.clip(vt + a * d, tmin, tmax).astype(np.float32), grid)
            trial.append(objective(m, tr)[0])
        # the parabola through (0, J), (a0, j1), (2 a0, j2): vertex at
        # a0 (3J + j2 - 4 j1) / (2 (j2 - 2 j1 + J)) when convex
        j1, j2 = trial
        den = j2 - 2 * j1 + J
        alpha = a0 * 0.5 * (3 * J + j2 - 4 * j1) / den if den > 0 else (2 * a0 if j2 < j1 else a0)
        alpha = float(np.clip(alpha, 0.25 * a0, 4 * a0))
        vt = _refill(np.clip(vt + alpha * d, tmin, tmax).astype(np.float32), grid)
        objs.append(J)
        if log:
            log(f"  iter {it}: objective {J:.5e}, gradient {gms[-1]:.2f} ms/shot, step {alpha:.3e}")
        if it + 1 in keep:
            kept[it + 1] = physical(vt)
    b.close()
    tr.close()
    return physical(vt), objs, float(np.median(gms)) if gms else 0.0, pts, kept


def upsample(vc, shape):
    """Coarse physical model -> fine: each fine cell takes the coarse cell
    nearest to it (fine (2i, 2k + 1) is coarse (i, k))."""
    ix = np.minimum((np.arange(shape[0]) + 1) // 2, vc.shape[0] - 1)
    iz = np.clip((np.arange(shape[1]) - OZ + 1) // 2, 0, vc.shape[1] - 1)
    return np.ascontiguousarray(vc[np.ix_(ix, iz)])


def fine_misfit(case, grid_f, v, dobs_f):
    """The judge: the fine propagator's data misfit of a physical model."""
    b = _batch(grid_f, capi.SNAP_NONE, capi.ORDER_10)
    b.bind_model(0, 0, expand(v, grid_f))
    d = b.forward()
    b.close()
    return float(np.sum((np.asarray(dobs_f, np.float64).ravel() - d) ** 2))
