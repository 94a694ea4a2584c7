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

typedef struct swfwi_solver_s *swfwi_solver;       /* Damp4t10d: geometry + bound model */
typedef struct swfwi_inversion_s *swfwi_inversion; /* FwiFramework/EssFwiFramework state */
typedef struct swfwi_enkf_s *swfwi_enkf;           /* EnkfAnalyze state */
typedef struct swfwi_enfwi_s *swfwi_enfwi;         /* ensemble of EssFwiFramework drivers */

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
 * recur (on by default inside the inversion drivers' epochs). */
int swfwi_solver_set_graphs(swfwi_solver h, int on);
/* Stencil tile width along x: 8, 16 or 32, or 0 for the measured rule
 * by batch size (the default). */
int swfwi_solver_set_tile_x(swfwi_solver h, int tile_x);
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

#ifdef __cplusplus
}
#endif

#endif /* SWFWI_HIP_H_ */
