/*
 * swfwi_hip.h -- C ABI of the MI355X acoustic FD propagation and
 * adjoint-gradient library (libswfwi_hip.so).
 *
 * Plain C types only.  Every entry point except the two per-step drop-in
 * kernels returns a status code (SWFWI_OK on success); swfwi_last_error()
 * gives the message of the last failure on the calling thread.
 *
 * Grid conventions follow the reference (damp4t10d.cpp:138-151,174): a
 * physical model of nz x nx cells is expanded to nzpad = nz+nb+12 by
 * nxpad = nx+2nb+12 with the sponge on the left, right and bottom, stored
 * index ix*nzpad + iz (z contiguous), in transformed units
 * vel_t = dx^2/(v^2 dt^2).  Source/receiver positions are physical cell
 * indices (ShotPosition::getx/getz, shot-position.cpp:11-37).  Seismograms
 * are time-major [nt][ng] (Damp4t10d::recordSeis, damp4t10d.cpp:163-180).
 */
#ifndef SWFWI_HIP_H_
#define SWFWI_HIP_H_

#ifdef __cplusplus
extern "C" {
#endif

#define SWFWI_OK 0
#define SWFWI_ERR_ARG 1     /* invalid argument or geometry */
#define SWFWI_ERR_STATE 2   /* call out of order (e.g. no velocity bound) */
#define SWFWI_ERR_HIP 3     /* HIP runtime failure */
#define SWFWI_ERR_NUMERIC 4 /* numerical failure (eigen-solve did not converge) */
#define SWFWI_ERR_RANGE 5   /* size beyond a supported maximum */

#define SWFWI_PREC_EXACT 0 /* the reference's FP32/FP64 expression trees, bit-for-bit */
#define SWFWI_PREC_FP32 1  /* pure FP32 with fused multiply-adds */

#define SWFWI_SNAP_F32 0  /* wavefield checkpoints stored in FP32 */
#define SWFWI_SNAP_BF16 1 /* ... in BF16 with a per-snapshot max-abs scale */
#define SWFWI_SNAP_F16 2  /* ... in FP16 with a per-snapshot max-abs scale */
#define SWFWI_SNAP_NONE 3 /* no checkpoints (forward runs only) */

#define SWFWI_ORDER_10 10 /* the reference's damped 10th-order-in-space, 4th-order-in-time step */
#define SWFWI_ORDER_4 4   /* damped 4th-order-in-space, 2nd-order-in-time step for coarse grids */

typedef struct swfwi_solver_s *swfwi_solver;       /* Damp4t10d: geometry + bound model */
typedef struct swfwi_inversion_s *swfwi_inversion; /* FwiFramework/EssFwiFramework state */
typedef struct swfwi_enkf_s *swfwi_enkf;           /* EnkfAnalyze state */
typedef struct swfwi_enfwi_s *swfwi_enfwi;         /* ensemble of EssFwiFramework drivers */
typedef struct swfwi_batch_s *swfwi_batch;         /* mixed-size batch with reduced-precision checkpoints */
typedef struct swfwi_net_s *swfwi_net;             /* learned coarse-grid correction network */

const char *swfwi_last_error(void);
const char *swfwi_version(void);
/* Number of visible HIP devices (0 when none); never initialises a context. */
int swfwi_device_count(int *count);
/* Practical bandwidth ceiling: a plain float4 device copy of n floats,
 * `iters` launches; *gbs = (read + write bytes) / time.  No counterpart in
 * the reference: it calibrates the roofline and the HBM byte counters. */
int swfwi_copy_bandwidth(long long n, int iters, double *gbs);
/* Microseconds per launch of `reps` back-to-back near-empty kernels of
 * `blocks` blocks, launched on a stream or replayed from a captured graph. */
int swfwi_launch_latency(int blocks, int reps, int graph, double *us);

/* ---- drop-in per-step kernels (host pointers, in place prev <- next) ----- */
/* Replaces fd4t10s_damp_zjh_2d_vtrans (fd4t10s-damp-zjh.h:11,
 * fd4t10s-damp-zjh.c:14-115): nx, nz are the padded dims, nb the damping width
 * (Damp4t10d passes bx0).  u2 receives the unscaled Laplacian over [5, n-5)
 * exactly as the reference's scratch does.  Aborts the process on a device
 * error, as the reference's error paths exit. */
void fd4t10s_damp_zjh_2d_vtrans(float *prev_wave, const float *curr_wave, const float *vel, float *u2, int nx,
                                int nz, int nb);
/* Replaces fd4t10s_zjh_2d_vtrans (fd4t10s-zjh.h:11, fd4t10s-zjh.c:10-56). */
void fd4t10s_zjh_2d_vtrans(float *prev_wave, const float *curr_wave, const float *vel, float *u2, int nx, int nz);
/* The same two kernels with a status return and a precision switch. */
int swfwi_step_damped(float *prev_wave, const float *curr_wave, const float *vel, float *u2, int nx, int nz, int nb,
                      int precision);
int swfwi_step_undamped(float *prev_wave, const float *curr_wave, const float *vel, float *u2, int nx, int nz,
                        int precision);

/* ---- host helpers of the Damp4t10d surface -------------------------------- */
/* Damp4t10d::expandDomain (damp4t10d.cpp:138-151): v [nx][nz] m/s ->
 * out [nxpad][nzpad] transformed. */
int swfwi_expand_domain(const float *v, int nx, int nz, int nb, float dx, float dt, float *out);
/* rickerWavelet (ricker-wavelet.cpp:15-21) */
int swfwi_ricker(float *wlt, int nt, float fm, float dt, float amp);
/* RandomCodes(seed).genPlus1Minus1(n) after `skip` earlier draws
 * (random-code.cpp:2-22). */
int swfwi_random_codes(int seed, int skip, int n, int *codes);

/* ---- device-resident propagator (Damp4t10d) ------------------------------ */
/* Damp4t10d(allSrcPos, allGeoPos, dt, dx, fm, nb, nt) (damp4t10d.cpp:367-376)
 * on the current HIP device, one stream per handle; handles are not
 * thread-safe, many may coexist.  nb must be >= 6. */
int swfwi_solver_create(swfwi_solver *out, int nx, int nz, int nb, float dt, float dx, float fm, int nt, int ns,
                        const int *sx, const int *sz, int ng, const int *gx, const int *gz, int precision);
int swfwi_solver_destroy(swfwi_solver h);
/* dims[0..5] = nxpad, nzpad, bx0, bz0, boundary floats per step, stencil tiles */
int swfwi_solver_dims(swfwi_solver h, int *dims);
int swfwi_solver_set_precision(swfwi_solver h, int precision);
/* Damp4t10d::bindVelocity (damp4t10d.cpp:159-161): padded transformed model */
int swfwi_solver_bind_velocity(swfwi_solver h, const float *vel_t);
int swfwi_solver_get_velocity(swfwi_solver h, float *vel_t);
/* Damp4t10d::writeBndry / readBndry for one step (damp4t10d.cpp:478-536):
 * the 6-wide boundary band of a padded field [nxpad][nzpad] to bndr[dims[4]]
 * and back (read_bndry updates only the band points of `field`). */
int swfwi_write_bndry(swfwi_solver h, const float *field, float *bndr);
int swfwi_read_bndry(swfwi_solver h, const float *bndr, float *field);
/* FwiForwardModeling (shot >= 0; encsrc[nt], damp4t10d.cpp:272-289) or
 * EssForwardModeling (shot < 0; encsrc[nt][ns], damp4t10d.cpp:291-307). */
int swfwi_forward_modeling(swfwi_solver h, const float *encsrc, int shot, float *dcal);
/* swfwi_forward_modeling without any device allocation.  No counterpart in
 * the reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as swfwi_forward_modeling; the wavefields, the
 * device seismogram and the source table live in a fixed per-GPU workspace
 * of 64 MiB held until the process ends: a geometry needing more than
 * 2*nxpad*nzpad + nt*ng + nt*ns floats (each region rounded up to 64 floats)
 * beyond that returns SWFWI_ERR_RANGE.  Every call on one GPU shares that
 * GPU's workspace: calls that overlap on one GPU race. */
int swfwi_forward_modeling_noalloc(swfwi_solver h, const float *encsrc, int shot, float *dcal);
/* calgradient (fwiframework.cpp:114-177 for shot >= 0,
 * essfwiframework.cpp:114-173 for shot < 0): vsrc [nt][ng] adjoint source,
 * grad [nxpad][nzpad] receives the raw imaging condition (before
 * scaleGradient/maskGradient). */
int swfwi_calgradient(swfwi_solver h, const float *encsrc, int shot, const float *vsrc, float *grad);
/* Observed data of every shot, one batched launch per time step (fm-damp's
 * shot loop, main-fm-damp.cpp:241-256): wlt [nt] -> dobs [ns][nt][ng]. */
int swfwi_forward_shots(swfwi_solver h, const float *wlt, float *dobs);
/* Grid-point updates (entry-weighted; the backward kernel counts its two
 * fields) and stencil launches issued by this handle so far, [0] forward
 * step kernel, [1] backward step kernel. */
int swfwi_solver_counters(swfwi_solver h, double *point_updates, long long *launches);
/* Capture time loops into HIP graphs and replay them when the same buffers
 * recur: 0 off, 1 batches of up to 8 entries (larger batches are not
 * launch-bound; on by default inside the inversion drivers' epochs), 2 every
 * batch size. */
int swfwi_solver_set_graphs(swfwi_solver h, int on);
/* Stencil tile width along x: 8, 16 or 32, or 0 for the measured rule
 * by batch size (the default). */
int swfwi_solver_set_tile_x(swfwi_solver h, int tile_x);
/* The tile tuner (1, the default) or the measured rule alone (0): with
 * tile_x 0, the first forward run of each batch size times every width on
 * scratch wavefields and keeps the rule's width unless another is faster by
 * 3% in both timed rounds.  Every width gives the same results bit for bit. */
int swfwi_solver_set_tuning(swfwi_solver h, int on);
/* Tile width in use for a kernel kind (0 forward, 1 backward) and batch
 * size, 0 when that batch size has not run yet. */
int swfwi_solver_get_tile_x(swfwi_solver h, int kind, int entries, int *tile_x);
/* Backward iterations as one fused block per tile (0), as separate source- and
 * receiver-field blocks (1), or by the measured rule (-1, the default: split
 * while the batch's wavefields fit the last-level cache, fused beyond).
 * Same results. */
int swfwi_solver_set_bwd_split(swfwi_solver h, int mode);
/* Record a HIP event pair around every forward / backward time loop on the
 * handle's stream (graph replays included); swfwi_solver_kernel_times waits
 * for them and returns the accumulated milliseconds and step-kernel launch
 * counts of the forward [0] and backward [1] loops. */
int swfwi_solver_set_timing(swfwi_solver h, int on);
int swfwi_solver_kernel_times(swfwi_solver h, double *ms, long long *launches);
/* Microseconds per launch of `reps` back-to-back step kernels (kind 0 forward,
 * 1 backward) of width tile_x (8, 16, 32) over `entries` scratch wavefields of
 * the bound model: the tile tuner's measurement, for sweeps.  Uses the
 * handle's workspace, so a following calgradient needs a new forward run. */
int swfwi_solver_time_steps(swfwi_solver h, int kind, int entries, int tile_x, int reps, float *us);

/* ---- inversion drivers (FwiFramework / EssFwiFramework) ------------------ */
/* wlt [nt], dobs [ns][nt][ng]; UpdateVelOp(vmin, vmax, dx, dt) and
 * UpdateSteplenOp(fm, velop, nita, maxdv) parameters. */
int swfwi_inversion_create(swfwi_inversion *out, swfwi_solver h, const float *wlt, const float *dobs, float vmin,
                           float vmax, int nita, float maxdv);
int swfwi_inversion_destroy(swfwi_inversion inv);
/* EssFwiFramework::epoch (essfwiframework.cpp:191-239) */
int swfwi_ess_epoch(swfwi_inversion inv, int iter);
/* FwiFramework::epoch (fwiframework.cpp:269-356), all shots batched; the
 * reference's step length is undefined behaviour, this uses the defined
 * two-point parabolic rule documented with the library. */
int swfwi_fwi_epoch(swfwi_inversion inv, int iter);
/* out[10] = obj1, steplen, alpha2, alpha3, obj2, obj3, updateobj, initobj,
 * trial forward runs, parabolic flag; ms[2] = gradient, line-search time. */
int swfwi_inversion_stats(swfwi_inversion inv, float *out, double *ms);
int swfwi_inversion_gradient(swfwi_inversion inv, float *grad);   /* scaled + masked */
int swfwi_inversion_direction(swfwi_inversion inv, float *dir);   /* CG direction */

/* ---- ensemble Kalman analysis (EnkfAnalyze) ------------------------------- */
/* EnkfAnalyze(fm, wlt, dobs, sigfac) (enkfanalyze.h:21-29) for `nmembers`
 * velocity members on the solver's geometry (members live on the device). */
int swfwi_enkf_create(swfwi_enkf *out, swfwi_solver h, const float *wlt, const float *dobs, float sigfac,
                      int nmembers);
int swfwi_enkf_destroy(swfwi_enkf e);
/* members [N][nxpad][nzpad] in transformed units (velSet) */
int swfwi_enkf_set_members(swfwi_enkf e, const float *vels);
int swfwi_enkf_get_members(swfwi_enkf e, float *vels);
/* EnkfAnalyze::pAnalyze (enkfanalyze.cpp:113-183, gain 343-495): new ENKF
 * codes, batched encoded forward runs of all members, FP64 analysis, member
 * update.  perturbation [N][nt][ng] FP64 may be given; when null it is drawn
 * N(0, sigma) from a counter-based generator seeded with `seed`. */
int swfwi_enkf_analyze(swfwi_enkf e, const double *perturbation, unsigned long long seed);
/* EnkfAnalyze::pCreateAMean (enkfanalyze.cpp:578-597): mean [nxpad][nzpad] */
int swfwi_enkf_mean(swfwi_enkf e, float *mean);
/* sigma of the perturbation, clip position, singular values [N] and the
 * N x N gain t4 (column major) of the last analysis; any pointer may be null */
int swfwi_enkf_info(swfwi_enkf e, double *sigma, int *clip, double *svals, double *gain);

/* ---- ensemble-Kalman FWI (the enfwi-damp driver) --------------------------- */
/* One EssFwiFramework + UpdateSteplenOp(nita, maxdv) + UpdateVelOp(vmin, vmax)
 * per member of `e` (main-enfwi-damp.cpp:509-516); the members' velocities are
 * the ones held by `e`, updated in place. */
int swfwi_enfwi_create(swfwi_enfwi *out, swfwi_solver h, swfwi_enkf e, float vmin, float vmax, int nita,
                       float maxdv);
int swfwi_enfwi_destroy(swfwi_enfwi f);
/* essfwis[i]->epoch(iter) for every member (main-enfwi-damp.cpp:563-567), all
 * members' forward runs, gradients and line-search trials batched */
int swfwi_enfwi_epoch(swfwi_enfwi f, int iter);
/* calobj of the ensemble-mean model (main-enfwi-damp.cpp:262-283,586-603);
 * mean [nxpad][nzpad] (may be null) receives the mean model */
int swfwi_enfwi_mean_objective(swfwi_enfwi f, float *mean, float *obj);
/* out[N][10] per member as swfwi_inversion_stats; ms[2] = gradient and
 * line-search time of the whole ensemble; counts[2] = batched trial rounds,
 * trial forward runs.  Any pointer may be null. */
int swfwi_enfwi_stats(swfwi_enfwi f, float *out, double *ms, int *counts);
/* scaled + masked gradients of the last epoch, [N][nxpad][nzpad] */
int swfwi_enfwi_gradients(swfwi_enfwi f, float *grads);

/* ---- mixed-size batched gradients with reduced-precision checkpoints ----- */
/* No single counterpart in the reference: calgradient (fwiframework.cpp:
 * 114-177) for many shots of many models on grids of different sizes, all
 * advancing in one launch per time step, with the source field kept as
 * checkpoints every 4 steps (in place of the boundary band of
 * damp4t10d.cpp:447-536 and the reconstruction it drives) and the imaging
 * condition formed on the checkpoint steps with weight 4 w(it).  The stencil
 * computes in FP32.
 * Group g: dims[9g..9g+8] = nx, nz, nb, nt, nshots, gx0, ng, gz, nmodels
 * (receivers at physical x = gx0 .. gx0+ng-1, depth gz); params[4g..4g+3] =
 * dx, dt, fm (Ricker wavelet of amplitude 1), sponge strength factor (1 for
 * the reference's profile, k for a grid stepping k times its dt); sx/sz: the shots' physical
 * positions, groups concatenated.  Entries (group, model, shot) are ordered
 * group by group, model by model, shot by shot in every array below. */
int swfwi_batch_create(swfwi_batch *out, int ngroups, const int *dims, const float *params, const int *sx,
                       const int *sz, int snap_prec, int order);
int swfwi_batch_destroy(swfwi_batch b);
/* BF16 <-> FP16 at any time; FP32 only for a batch created in FP32 */
int swfwi_batch_set_snapshot_precision(swfwi_batch b, int snap_prec);
/* Capture each time loop into a HIP graph on first use (on by default). */
int swfwi_batch_set_graphs(swfwi_batch b, int on);
/* The tuner (1, the default) or the rules alone (0): the first forward run
 * times the step kernels at tile widths 16, 24 and 32 and both block orders
 * over every entry, and keeps the rules' choice (width 24 from 8000 tiles of
 * 64 x 16, else 16; order 1) unless another is faster by 3% in both timed
 * rounds.  swfwi_batch_tile_x receives the width in use. */
int swfwi_batch_set_tuning(swfwi_batch b, int on);
/* A fixed tile width, 16, 24 or 32, or 0 for the rule and the tuner. */
int swfwi_batch_set_tile_x(swfwi_batch b, int tile_x);
/* Block order of the step launches: 0 entry by entry, 1 the shots of each
 * model tile by tile (a model tile is read once from HBM and then from L2),
 * or -1 for the rule (1) and the tuner, which times both orders with the
 * widths.  Results are the same bit for bit in every order. */
int swfwi_batch_set_block_order(swfwi_batch b, int order);
int swfwi_batch_block_order(swfwi_batch b, int *order);
int swfwi_batch_tile_x(swfwi_batch b, int *tile_x);
/* model `model` of group `group`: padded transformed velocity (Damp4t10d) */
int swfwi_batch_bind_model(swfwi_batch b, int group, int model, const float *vel_t);
/* Forward run of every entry: seismograms into dcal[entry][nt][ng] (may be
 * null) and the checkpoints kept on the device. */
int swfwi_batch_forward(swfwi_batch b, float *dcal);
/* Adjoint sources vsrc = transVsrc(dobs - dcal) per entry on the device
 * (fwiframework.cpp:75-105,293-302 without the mute), from the last forward
 * run; vsrc (may be null) receives them. */
int swfwi_batch_adjoint_source(swfwi_batch b, const float *dobs, float *vsrc);
/* Backward run from the adjoint sources vsrc[entry][nt][ng] (null: the
 * ones swfwi_batch_adjoint_source left on the device): raw imaging condition
 * per entry into grads[entry][nxpad][nzpad] and per model summed over its
 * shots into model_grads[model][nxpad][nzpad] (either may be null). */
int swfwi_batch_backward(swfwi_batch b, const float *vsrc, float *grads, float *model_grads);
/* out[6] = entries, forward and backward point updates per run, checkpoint
 * bytes, milliseconds of the last forward and backward run. */
int swfwi_batch_info(swfwi_batch b, double *out);
/* Field taps of the forward runs: at each of the n listed steps, the
 * restriction of every entry's current field to the 2x-coarser grid is kept:
 * physical fine cell (2i, 2k + oz) -> coarse cell (i, k) of the grid of
 * nx' = (nx+1)/2, nz' = (nz+1-oz)/2, nb' = nb/2 (nb must be even; oz = 1
 * puts both grids' zero top rows at the same depth); swfwi_batch_taps copies
 * them out as [entry][n][nxpad'][nzpad']. */
int swfwi_batch_set_taps(swfwi_batch b, int n, const int *steps, int oz);
int swfwi_batch_taps(swfwi_batch b, float *out);
/* Steps it0 .. it0+nsteps-1 of every entry from the state prev (field at
 * it0-1) and cur (at it0), [entry][nxpad][nzpad] each; the state after them
 * (prev_out at it0+nsteps-1, cur_out at it0+nsteps; either may be null). */
int swfwi_batch_step_from(swfwi_batch b, const float *prev, const float *cur, int it0, int nsteps,
                          float *prev_out, float *cur_out);
/* The learned correction p <- p + net(p) on both time levels of every entry
 * every 4 forward steps, over the physical interior (null: off).
 * Single-group batches.  The batch keeps a pointer to the network: destroy
 * the network only after the batch, or after setting null here. */
int swfwi_batch_set_correction(swfwi_batch b, swfwi_net net);

/* ---- learned coarse-grid correction network --------------------------------
 * No counterpart in the reference.  A 3x3 convolution network 2 -> 16 -> 16
 * -> 1 with ReLU between and no biases, y = net(p) added to the field p; its
 * inputs are p and r p, r = 1 / vel_t (the local Courant number squared,
 * given as the rv stack beside the fields).  Weights (2736 floats):
 * W1[t][k][c] (9 x 2 x 16), W2[t][ci][co] (9 x 16 x 16), W3[t][ci] (9 x 16),
 * tap t = (dx + 1) * 3 + (dz + 1) for the neighbour (x + dx, z + dz).  Every
 * layer and the weight-gradient products run on BF16 MFMA with FP32
 * accumulation (inputs, activations and weights rounded to BF16).  Field stacks are [entries][nx][nz] (z contiguous);
 * region[4] = x0, x1, z0, z1: the outputs and the loss cover
 * [x0, x1) x [z0, z1), which must lie 5 cells inside the array. */
int swfwi_net_create(swfwi_net *out);
int swfwi_net_destroy(swfwi_net n);
int swfwi_net_set_weights(swfwi_net n, const float *w);
int swfwi_net_get_weights(swfwi_net n, float *w);
/* y = net(fields) over the region (zero elsewhere) */
int swfwi_net_apply(swfwi_net n, const float *fields, const float *rv, int nx, int nz, int entries, const int *region,
                    float *y);
/* Milliseconds per launch of the network's forward kernel over the fields
 * (uploaded once), averaged over `reps` back-to-back launches.  No
 * counterpart in the reference: it measures the kernel for its roofline. */
int swfwi_net_time_apply(swfwi_net n, const float *fields, const float *rv, int nx, int nz, int entries,
                         const int *region, int reps, double *ms);
/* loss = mean over the region of (in + net(in) - target)^2 and its gradient
 * with respect to the weights (grad may be null) */
int swfwi_net_loss_grad(swfwi_net n, const float *in, const float *rv, const float *target, int nx, int nz,
                        int entries, const int *region, float *grad, double *loss);
/* Training set kept on the device: nsamples inputs, their rv and targets
 * [n][nx][nz] each. */
int swfwi_net_set_data(swfwi_net n, const float *in, const float *rv, const float *target, int nx, int nz,
                       int nsamples, const int *region);
/* `steps` Adam steps (beta 0.9 / 0.999) on minibatches of `batch` samples in a
 * seeded shuffled order; losses[steps] (may be null) = loss before each step. */
int swfwi_net_train(swfwi_net n, int batch, int steps, float lr, unsigned long long seed, double *losses);

#ifdef __cplusplus
}
#endif

#endif /* SWFWI_HIP_H_ */
