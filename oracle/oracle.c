/*
 * oracle.c -- CPU restatement of the swfwi acoustic FD propagation and
 * adjoint-gradient path, the reference results of the parity tests.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is built into
 * oracle/liboracle.so and loaded by tests/, by __graft_entry__.smoke() and by
 * the CPU-baseline leg of bench.py, where it is the thing the GPU results are
 * checked against (or timed beside).  The product library never links or
 * calls it.
 *
 * Every function restates one piece of the reference algorithm and cites the
 * reference file:line it follows.  The floating-point expression trees are
 * kept exactly as the reference writes them (float vs double sub-expressions,
 * evaluation order), because the reference's arithmetic is the spec: compile
 * with -O2 -ffp-contract=off and no -march (SSE2, no FMA), as the reference's
 * gnu build does.
 *
 * Reductions: the reference sums objectives and CG dot products serially in
 * FP32.  `orc_sum_mode` selects that (0, the reference) or FP64 accumulation
 * (1, what the GPU path does) so the tests can pin the downstream control
 * flow against either.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EXFD 6 /* stencil halo, damp4t10d.h:69 (EXFDBNDRYLEN) */

static int orc_sum_mode = 0;
void orc_set_sum_mode(int mode) { orc_sum_mode = mode; }
int orc_get_sum_mode(void) { return orc_sum_mode; }

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* 1. The two stencil kernels.                                              */
/* ------------------------------------------------------------------------ */

/* Zhang Jinhai 10th-order coefficients, fd4t10s-damp-zjh.c:24-29 and
 * fd4t10s-zjh.c:16-21, stored as float exactly as the reference does. */
static const float ZJH[6] = {+1.53400796f, +1.78858721f, -0.31660756f,
                             +0.07612173f, -0.01626042f, +0.00216736f};

/* u2 at one point: the unscaled 10th-order Laplacian.  The first term is a
 * double product (-4.0*a0*p), each ring a float product, the sum in double,
 * stored as float (fd4t10s-damp-zjh.c:44-54, fd4t10s-zjh.c:28-38). */
static inline float orc_u2_point(const float *c, int i, int nz) {
  double acc = -4.0 * ZJH[0] * c[i];
  for (int k = 1; k <= 5; k++) {
    float ring = c[i - k] + c[i + k];
    ring = ring + c[i - k * nz];
    ring = ring + c[i + k * nz];
    acc = acc + (float)(ZJH[k] * ring);
  }
  return (float)acc;
}

static void orc_u2_pass(const float *cur, float *u2, int nx, int nz, int parallel) {
  const int d = EXFD;
  if (parallel) {
#pragma omp parallel for schedule(static)
    for (int ix = d - 1; ix < nx - (d - 1); ix++)
      for (int iz = d - 1; iz < nz - (d - 1); iz++) u2[ix * nz + iz] = orc_u2_point(cur, ix * nz + iz, nz);
  } else {
    for (int ix = d - 1; ix < nx - (d - 1); ix++)
      for (int iz = d - 1; iz < nz - (d - 1); iz++) u2[ix * nz + iz] = orc_u2_point(cur, ix * nz + iz, nz);
  }
}

/* The 4th-order time correction term r^4/12 * Lap5(u2), all float
 * (fd4t10s-damp-zjh.c:87-88). */
static inline float orc_t4_term(const float *u2, int i, int nz, float rv) {
  float lap = u2[i - 1] + u2[i + 1];
  lap = lap + u2[i - nz];
  lap = lap + u2[i + nz];
  lap = lap - 4 * u2[i];
  float c = 1.0f / 12 * rv;
  c = c * rv;
  return c * lap;
}

/* Damping profile of the damped kernel at one point (fd4t10s-damp-zjh.c:64-80):
 * the z (bottom) branch is tested last and wins in the corners. */
static inline float orc_delta(int ix, int iz, int nx, int nz, int nb) {
  const float max_delta = 0.05f;
  const int bx = nb, bz = nb;
  float dist = 0;
  if (ix < bx) dist = (float)(bx - ix) / bx;
  if (ix >= nx - bx) dist = (float)(ix - (nx - bx) + 1) / bx;
  if (iz >= nz - bz) dist = (float)(iz - (nz - bz) + 1) / bz;
  return max_delta * dist * dist;
}

/* fd4t10s_damp_zjh_2d_vtrans (fd4t10s-damp-zjh.c:14-115).  `parallel_u2`
 * selects the threading of the u2 pass: 0 = serial, as the reference build
 * runs it (USE_OPENMP is never defined, fd4t10s-damp-zjh.c:32-34); 1 = the
 * all-parallel variant.  Results are identical either way. */
void orc_fd4t10s_damp(float *prev, const float *cur, const float *vel, float *u2, int nx, int nz, int nb,
                      int parallel_u2) {
  const int d = EXFD;
  orc_u2_pass(cur, u2, nx, nz, parallel_u2);
#pragma omp parallel for schedule(static)
  for (int ix = d; ix < nx - d; ix++) {
    for (int iz = d; iz < nz - d; iz++) {
      int i = ix * nz + iz;
      float delta = orc_delta(ix, iz, nx, nz, nb);
      float rv = 1.0f / vel[i];
      double acc = (2. - 2 * delta + delta * delta) * cur[i];
      acc = acc - (float)((1 - 2 * delta) * prev[i]);
      acc = acc + (float)(rv * u2[i]);
      acc = acc + orc_t4_term(u2, i, nz, rv);
      prev[i] = (float)acc;
    }
  }
}

/* fd4t10s_zjh_2d_vtrans (fd4t10s-zjh.c:10-56): 2.*p is a double term. */
void orc_fd4t10s(float *prev, const float *cur, const float *vel, float *u2, int nx, int nz) {
  const int d = EXFD;
  orc_u2_pass(cur, u2, nx, nz, 1);
#pragma omp parallel for schedule(static)
  for (int ix = d; ix < nx - d; ix++) {
    for (int iz = d; iz < nz - d; iz++) {
      int i = ix * nz + iz;
      float rv = 1.0f / vel[i];
      double acc = 2. * cur[i] - prev[i];
      acc = acc + (float)(rv * u2[i]);
      acc = acc + orc_t4_term(u2, i, nz, rv);
      prev[i] = (float)acc;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* 2. Domain expansion and velocity transforms (damp4t10d.cpp:47-151).      */
/* ------------------------------------------------------------------------ */

/* expandBndry (damp4t10d.cpp:96-124) + transvel (:126-130) +
 * expandForStencil/fillForStencil (:47-94).  v is [nx][nz] physical m/s;
 * out is [nx+2nb+12][nz+nb+12] in transformed units dx^2/(v^2 dt^2). */
void orc_expand_domain(const float *v, int nx, int nz, int nb, float dx, float dt, float *out) {
  const int h = EXFD;
  const int nxb = nx + 2 * nb, nzb = nz + nb;
  float *b = (float *)calloc((size_t)nxb * nzb, sizeof(float));
  for (int ix = 0; ix < nx; ix++)
    for (int iz = 0; iz < nz; iz++) b[(nb + ix) * nzb + iz] = v[ix * nz + iz];
  for (int ix = 0; ix < nb; ix++)
    for (int iz = 0; iz < nz; iz++) {
      b[ix * nzb + iz] = v[iz];
      b[(nb + nx + ix) * nzb + iz] = v[(nx - 1) * nz + iz];
    }
  for (int ix = 0; ix < nxb; ix++)
    for (int iz = 0; iz < nb; iz++) b[ix * nzb + (nz + iz)] = b[ix * nzb + (nz - 1)];
  for (size_t i = 0; i < (size_t)nxb * nzb; i++) b[i] = (dx * dx) / (b[i] * b[i] * dt * dt);

  const int nxp = nxb + 2 * h, nzp = nzb + 2 * h;
  for (int ix = h; ix < nxb + h; ix++)
    for (int iz = h; iz < nzb + h; iz++) out[ix * nzp + iz] = b[(ix - h) * nzb + (iz - h)];
  for (int ix = h; ix < nxb + h; ix++) {
    for (int iz = 0; iz < h; iz++) out[ix * nzp + iz] = out[ix * nzp + h];
    for (int iz = nzb + h; iz < nzp; iz++) out[ix * nzp + iz] = out[ix * nzp + (nzb + h - 1)];
  }
  for (int iz = 0; iz < nzp; iz++) {
    for (int ix = 0; ix < h; ix++) out[ix * nzp + iz] = out[h * nzp + iz];
    for (int ix = nxb + h; ix < nxp; ix++) out[ix * nzp + iz] = out[(nxb + h - 1) * nzp + iz];
  }
  free(b);
}

/* recoverVel (damp4t10d.cpp:132-136) */
void orc_recover_vel(float *vel, long n, float dx, float dt) {
  for (long i = 0; i < n; i++) vel[i] = sqrtf(dx * dx / (dt * dt * vel[i]));
}

/* velTrans / velRecover templates (common.h:21-30) -- note the different
 * multiplication order from transvel/recoverVel. */
float orc_vel_trans(float vel, float dx, float dt) { return dx * dx / (dt * dt * vel * vel); }
float orc_vel_recover(float vel, float dx, float dt) {
  float t = dx * dx / (dt * dt * vel);
  return sqrtf(t);
}

/* ------------------------------------------------------------------------ */
/* 3. The propagator state (Damp4t10d members, damp4t10d.h:71-85).          */
/* ------------------------------------------------------------------------ */

typedef struct {
  int nx, nz;                 /* padded dims: vel->nx, vel->nz */
  int bx0, bxn, bz0, bzn;     /* damp4t10d.cpp:372-373 */
  float dt, dx, fm;
  int nt;
  int ns;                     /* all sources */
  const int *sx, *sz;         /* physical cell coords (ShotPosition getx/getz) */
  int ng;
  const int *gx, *gz;
  int parallel_u2;            /* threading of the damped kernel's u2 pass */
} orc_geom;

static float *orc_u2_scratch(const orc_geom *g) {
  return (float *)calloc((size_t)g->nx * g->nz, sizeof(float));
}

static void orc_step_forward(const orc_geom *g, const float *vel, float *p0, const float *p1, float *u2) {
  orc_fd4t10s_damp(p0, p1, vel, u2, g->nx, g->nz, g->bx0, g->parallel_u2); /* damp4t10d.cpp:153-157 */
}
static void orc_step_backward(const orc_geom *g, const float *vel, float *p0, const float *p1, float *u2) {
  orc_fd4t10s(p0, p1, vel, u2, g->nx, g->nz); /* damp4t10d.cpp:188-191 */
}

/* manipSource (damp4t10d.cpp:204-214): sources s in [lo, hi) of the position
 * list, amplitude src[s - lo], sign +1 add / -1 subtract, applied in order. */
static void orc_manip(const orc_geom *g, float *p, const float *src, const int *px, const int *pz, int lo, int hi,
                      int sign) {
  for (int s = lo; s < hi; s++) {
    int ix = px[s] + g->bx0, iz = pz[s] + g->bz0;
    long i = (long)ix * g->nz + iz;
    if (sign > 0)
      p[i] = p[i] + src[s - lo];
    else
      p[i] = p[i] - src[s - lo];
  }
}

/* recordSeis (damp4t10d.cpp:163-180) */
static void orc_record(const orc_geom *g, float *seis, const float *p) {
  for (int r = 0; r < g->ng; r++) seis[r] = p[(long)(g->gx[r] + g->bx0) * g->nz + (g->gz[r] + g->bz0)];
}

/* initBndryVector (damp4t10d.cpp:447-476): size of one step's band. */
int orc_bndr_size(const orc_geom *g) {
  const int w = 6;
  int nxb = g->nx - (g->bx0 - w + g->bxn - w);
  int nzb = g->nz - g->bz0 - g->bzn;
  return w * (nxb + 2 * nzb);
}

/* writeBndry / readBndry (damp4t10d.cpp:478-536) */
static void orc_rw_bndr(const orc_geom *g, float *bndr, float *p, int write) {
  const int w = 6, nzp = g->nz, nxp = g->nx;
  int nxb = nxp - (g->bx0 - w + g->bxn - w);
  int nzb = nzp - g->bz0 - g->bzn;
  for (int ix = 0; ix < nxb; ix++)
    for (int iz = 0; iz < w; iz++) {
      long pi = (long)(ix + g->bx0 - w) * nzp + (nzp - g->bzn + iz);
      int bi = iz + w * ix;
      if (write) bndr[bi] = p[pi]; else p[pi] = bndr[bi];
    }
  for (int iz = 0; iz < nzb; iz++)
    for (int ix = 0; ix < w; ix++) {
      long pl = (long)(g->bx0 - w + ix) * nzp + (g->bz0 + iz);
      long pr = (long)(nxp - g->bxn + ix) * nzp + (g->bz0 + iz);
      int bl = w * nxb + iz + nzb * ix, br = w * nxb + iz + nzb * (ix + w);
      if (write) { bndr[bl] = p[pl]; bndr[br] = p[pr]; }
      else { p[pl] = bndr[bl]; p[pr] = bndr[br]; }
    }
}

/* FwiForwardModeling (shot >= 0: source `shot` alone, amplitude encsrc[it],
 * damp4t10d.cpp:272-289) / EssForwardModeling (shot < 0: all ns sources,
 * amplitudes encsrc[it*ns + s], damp4t10d.cpp:291-307).  dcal is [nt][ng]. */
void orc_forward_modeling(const orc_geom *g, const float *vel, const float *encsrc, int shot, float *dcal) {
  size_t n = (size_t)g->nx * g->nz;
  float *p0 = (float *)calloc(n, sizeof(float)), *p1 = (float *)calloc(n, sizeof(float));
  float *u2 = orc_u2_scratch(g);
  for (int it = 0; it < g->nt; it++) {
    if (shot >= 0)
      orc_manip(g, p1, &encsrc[it], g->sx, g->sz, shot, shot + 1, +1);
    else
      orc_manip(g, p1, &encsrc[(size_t)it * g->ns], g->sx, g->sz, 0, g->ns, +1);
    orc_step_forward(g, vel, p0, p1, u2);
    float *t = p0; p0 = p1; p1 = t;
    orc_record(g, &dcal[(size_t)it * g->ng], p0);
  }
  free(p0); free(p1); free(u2);
}

/* The imaging-condition weight of backward step `it` (fwiframework.cpp:164-175):
 * float dt*it compared with double constants.  Returns 0 for "break". */
int orc_xcorr_weight(float dt, int it, float *w) {
  if (dt * it > 0.4) { *w = 1.0; return 1; }
  if (dt * it > 0.3) { *w = (dt * it - 0.3) / 0.1; return 1; }
  return 0;
}

/* calgradient (fwiframework.cpp:114-177 for shot >= 0, one source with
 * amplitude encsrc[it]; essfwiframework.cpp:114-173 for shot < 0, all sources
 * with encsrc[it*ns + s]).  vsrc is [nt][ng]; g0 (padded grid) is
 * accumulated into (the caller zeroes it, as the reference's g1). */
void orc_calgradient(const orc_geom *g, const float *vel, const float *encsrc, int shot, const float *vsrc, float *g0) {
  size_t n = (size_t)g->nx * g->nz;
  int bs = orc_bndr_size(g);
  float *bndr = (float *)calloc((size_t)g->nt * bs, sizeof(float));
  float *sp0 = (float *)calloc(n, sizeof(float)), *sp1 = (float *)calloc(n, sizeof(float));
  float *gp0 = (float *)calloc(n, sizeof(float)), *gp1 = (float *)calloc(n, sizeof(float));
  float *u2 = orc_u2_scratch(g);
  int lo = shot >= 0 ? shot : 0, hi = shot >= 0 ? shot + 1 : g->ns;

  for (int it = 0; it < g->nt; it++) {
    const float *s = shot >= 0 ? &encsrc[it] : &encsrc[(size_t)it * g->ns];
    orc_manip(g, sp1, s, g->sx, g->sz, lo, hi, +1);
    orc_step_forward(g, vel, sp0, sp1, u2);
    float *t = sp0; sp0 = sp1; sp1 = t;
    orc_rw_bndr(g, &bndr[(size_t)it * bs], sp0, 1);
  }
  for (int it = g->nt - 1; it >= 0; it--) {
    orc_rw_bndr(g, &bndr[(size_t)it * bs], sp0, 0);
    float *t = sp0; sp0 = sp1; sp1 = t;
    orc_step_backward(g, vel, sp0, sp1, u2);
    const float *s = shot >= 0 ? &encsrc[it] : &encsrc[(size_t)it * g->ns];
    orc_manip(g, sp0, s, g->sx, g->sz, lo, hi, -1);

    orc_manip(g, gp1, &vsrc[(size_t)it * g->ng], g->gx, g->gz, 0, g->ng, +1);
    orc_step_forward(g, vel, gp0, gp1, u2);
    t = gp0; gp0 = gp1; gp1 = t;

    float w;
    if (!orc_xcorr_weight(g->dt, it, &w)) break;
    for (size_t i = 0; i < n; i++) g0[i] -= sp0[i] * gp0[i] * w; /* cross_correlation, fwiframework.cpp:107-112 */
  }
  free(bndr); free(sp0); free(sp1); free(gp0); free(gp1); free(u2);
}

/* ------------------------------------------------------------------------ */
/* 4. Data-side helpers.                                                    */
/* ------------------------------------------------------------------------ */

/* second_order_virtual_source_forth_accuracy + transVsrc
 * (fwiframework.cpp:75-105): per trace, FP64 5-point 2nd derivative in t,
 * first two and last two samples zeroed.  vsrc is [nt][ng]. */
void orc_trans_vsrc(float *vsrc, int nt, int ng) {
  float *tr = (float *)malloc(sizeof(float) * nt);
  for (int r = 0; r < ng; r++) {
    for (int t = 0; t < nt; t++) tr[t] = vsrc[(size_t)t * ng + r];
    for (int i = 0; i < nt; i++) {
      float v;
      if (i <= 1 || i == nt - 1 || i == nt - 2)
        v = 0.0f;
      else
        v = -1. / 12 * tr[i - 2] + 4. / 3 * tr[i - 1] - 2.5 * tr[i] + 4. / 3 * tr[i + 1] - 1. / 12 * tr[i + 2];
      vsrc[(size_t)i * ng + r] = v;
    }
  }
  free(tr);
}

/* The direct-arrival mute window of removeDirectArrival
 * (damp4t10d.cpp:321-365): per trace [start, end).  `src` selects the source
 * whose position is used (the ESS variant uses source 0, :396-399; the FWI
 * variant the shot's own, :390-394). */
void orc_mute_windows(const orc_geom *g, const float *vel, int src, int *start, int *end) {
  float t_width = 1.5 / g->fm;
  int half_len = t_width / g->dt;
  int sx = g->sx[src] + g->bx0;
  int sz = g->sz[src] + g->bz0;
  int gz0 = g->gz[0] + g->bz0;
  float vel_average = 0.0;
  int gmin = sz < gz0 ? sz : gz0, gmax = sz > gz0 ? sz : gz0;
  for (int i = 0; i < g->nx; i++)
    for (int k = gmin; k <= gmax; k++) vel_average += vel[(size_t)i * g->nz + k];
  vel_average /= g->nx * (gmax - gmin + 1);
  for (int r = 0; r < g->ng; r++) {
    int gx = g->gx[r] + g->bx0, gz = g->gz[r] + g->bz0;
    float dist = (gx - sx) * (gx - sx) + (gz - sz) * (gz - sz);
    int t = (int)sqrtf(dist * vel_average);
    start[r] = t;
    end[r] = (t + 2 * half_len) > g->nt ? g->nt : (t + 2 * half_len);
  }
}

void orc_remove_direct_arrival(const orc_geom *g, const float *vel, int src, float *data) {
  int *st = (int *)malloc(sizeof(int) * g->ng), *en = (int *)malloc(sizeof(int) * g->ng);
  orc_mute_windows(g, vel, src, st, en);
  for (int r = 0; r < g->ng; r++)
    for (int t = st[r]; t < en[r]; t++) data[(size_t)t * g->ng + r] = 0.f;
  free(st); free(en);
}

/* cal_objective (common.cpp:231-241): serial FP32 sum of squares, or FP64
 * accumulation under sum mode 1. */
float orc_cal_objective(const float *d, long n) {
  if (orc_sum_mode == 1) {
    double s = 0;
    for (long i = 0; i < n; i++) s += (double)d[i] * (double)d[i];
    return (float)s;
  }
  float obj = 0;
  for (long i = 0; i < n; i++) { float a = d[i]; obj += a * a; }
  return obj;
}

/* ------------------------------------------------------------------------ */
/* 5. Gradient post-processing and model update.                            */
/* ------------------------------------------------------------------------ */

/* scaleGradient (damp4t10d.cpp:309-319) */
void orc_scale_gradient(const orc_geom *g, float *grad) {
  int nz = g->nz - g->bz0 - g->bzn;
  for (int ix = g->bx0; ix < g->nx - g->bxn; ix++)
    for (int iz = 1; iz < nz; iz++) grad[(size_t)ix * g->nz + iz + g->bz0] *= sqrtf((float)iz);
}

/* maskGradient (damp4t10d.cpp:216-226) */
void orc_mask_gradient(const orc_geom *g, float *grad) {
  for (int ix = 0; ix < g->nx; ix++)
    for (int iz = 0; iz < g->nz; iz++)
      if (ix < g->bx0 || iz < g->bz0 || ix >= g->nx - g->bxn || iz >= g->nz - g->bzn) grad[(size_t)ix * g->nz + iz] = 0.f;
}

/* refillBoundary (damp4t10d.cpp:228-252) */
void orc_refill_boundary(const orc_geom *g, float *a) {
  int nzp = g->nz, nxp = g->nx;
  for (int ix = 0; ix < nxp; ix++) {
    for (int iz = 0; iz < g->bz0; iz++) a[(size_t)ix * nzp + iz] = a[(size_t)ix * nzp + g->bz0];
    for (int iz = nzp - g->bzn; iz < nzp; iz++) a[(size_t)ix * nzp + iz] = a[(size_t)ix * nzp + nzp - g->bzn - 1];
  }
  for (int ix = 0; ix < g->bx0; ix++)
    for (int iz = 0; iz < nzp; iz++) a[(size_t)ix * nzp + iz] = a[(size_t)g->bx0 * nzp + iz];
  for (int ix = nxp - g->bxn; ix < nxp; ix++)
    for (int iz = 0; iz < nzp; iz++) a[(size_t)ix * nzp + iz] = a[(size_t)(nxp - g->bxn - 1) * nzp + iz];
}

/* updateGrad, PR+ conjugate gradient (fwiframework.cpp:42-72). */
void orc_update_grad(float *pre, const float *cur, float *dir, long n, int iter) {
  if (iter == 0) {
    memcpy(dir, cur, sizeof(float) * n);
    memcpy(pre, cur, sizeof(float) * n);
    return;
  }
  float beta, a, b, c;
  if (orc_sum_mode == 1) {
    double da = 0, db = 0, dc = 0;
    for (long i = 0; i < n; i++) {
      da += (double)cur[i] * cur[i]; db += (double)cur[i] * pre[i]; dc += (double)pre[i] * pre[i];
    }
    a = (float)da; b = (float)db; c = (float)dc;
  } else {
    a = b = c = 0.0f;
    for (long i = 0; i < n; i++) { a += cur[i] * cur[i]; b += cur[i] * pre[i]; c += pre[i] * pre[i]; }
  }
  beta = (a - b) / c;
  if (beta < 0.0f) beta = 0.0f;
  for (long i = 0; i < n; i++) dir[i] = cur[i] + beta * dir[i];
  memcpy(pre, cur, sizeof(float) * n);
}

/* UpdateVelOp ctor bounds (updatevelop.cpp:11-21): transformed-unit clamp. */
void orc_vel_bounds(float vmin, float vmax, float dx, float dt, float *tmin, float *tmax) {
  *tmin = (dx / dt / vmax) * (dx / dt / vmax);
  *tmax = (dx / dt / vmin) * (dx / dt / vmin);
}

/* update_vel (updatevelop.cpp:11-21 in the listing: lines 11-21) */
void orc_update_vel(float *nv, const float *v, const float *grad, long n, float steplen, float tmin, float tmax) {
  for (long i = 0; i < n; i++) {
    nv[i] = v[i] + steplen * grad[i];
    if (nv[i] > tmax) nv[i] = tmax;
    if (nv[i] < tmin) nv[i] = tmin;
  }
}

/* calMaxAlpha2_3 (updatesteplenop.cpp:25-47) */
void orc_cal_max_alpha23(const float *vel, const float *grad, long n, float dt, float dx, float maxdv, float *a2,
                         float *a3) {
  float alpha2 = FLT_MAX;
  for (long i = 0; i < n; i++) {
    float tmpv = dx / (dt * sqrtf(vel[i]));
    tmpv -= maxdv;
    tmpv = (dx / (dt * tmpv)) * (dx / (dt * tmpv));
    if (fabsf(grad[i]) < 1e-10) continue;
    if (alpha2 > (tmpv - vel[i]) / fabsf(grad[i])) alpha2 = (tmpv - vel[i]) / fabsf(grad[i]);
  }
  *a2 = alpha2;
  *a3 = 2 * alpha2;
}

/* calParabolaVertex + parabolaVertex (parabola-vertex.cpp:12-33) */
void orc_parabola_vertex(float x1, float y1, float x2, float y2, float x3, float y3, float max_alpha3, float *xv,
                         float *yv) {
  (void)max_alpha3;
  double k2 = (y3 - y2) / (x3 - x2);
  double k1 = (y2 - y1) / (x2 - x1);
  double denom = (x1 - x2) * (x1 - x3) * (x2 - x3);
  double A = (x3 * (y2 - y1) + x2 * (y1 - y3) + x1 * (y3 - y2)) / denom;
  double B = (x3 * x3 * (y1 - y2) + x2 * x2 * (y3 - y1) + x1 * x1 * (y2 - y3)) / denom;
  double C = (x2 * x3 * (x2 - x3) * y1 + x3 * x1 * (x3 - x1) * y2 + x1 * x2 * (x1 - x2) * y3) / denom;
  *xv = -B / (2 * A);
  *yv = C - B * B / (4 * A);
  double m = fabs(k2) > fabs(k1) ? fabs(k2) : fabs(k1);
  if (fabs(k2 - k1) < 0.001 * m || isnan(-*xv)) { *xv = x3; *yv = y3; }
}

/* ------------------------------------------------------------------------ */
/* 6. Wavelet, random codes, encoder.                                       */
/* ------------------------------------------------------------------------ */

/* rickerWavelet (ricker-wavelet.cpp:15-21) */
void orc_ricker(float *wlt, int nt, float fm, float dt, float amp) {
  for (int it = 0; it < nt; it++) {
    float tmp = M_PI * fm * (it * dt - 1.0 / fm);
    tmp *= tmp;
    wlt[it] = amp * (1.0 - 2.0 * tmp) * expf(-tmp);
  }
}

/* RandomCodes (random-code.cpp:2-22): minstd_rand (x <- 48271 x mod 2^31-1)
 * mapped to {0,1} by uniform_int's bucket (x-1)/1073741823, then 2u-1. */
typedef struct { uint64_t x; } orc_rng;
void orc_rng_seed(orc_rng *r, int seed) {
  uint64_t s = (uint64_t)(uint32_t)seed % 2147483647u;
  r->x = s == 0 ? 1 : s;
}
uint32_t orc_minstd_next(orc_rng *r) {
  r->x = (r->x * 48271u) % 2147483647u;
  return (uint32_t)r->x;
}
void orc_gen_codes(orc_rng *r, int n, int *codes) {
  for (int i = 0; i < n; i++) {
    uint32_t x = orc_minstd_next(r);
    codes[i] = (int)((x - 1u) / 1073741823u) * 2 - 1;
  }
}

/* Encoder::encodeSource / encodeObsData (encoder.cpp:14-42) */
void orc_encode_source(const int *code, int ns, const float *wlt, int nt, float *enc) {
  for (int it = 0; it < nt; it++)
    for (int is = 0; is < ns; is++) enc[(size_t)it * ns + is] = wlt[it] * code[is];
}
void orc_encode_obs(const int *code, int ns, const float *dobs, int nt, int ng, float *enc) {
  size_t m = (size_t)nt * ng;
  memset(enc, 0, sizeof(float) * m);
  for (int is = 0; is < ns; is++)
    for (size_t j = 0; j < m; j++) enc[j] += dobs[is * m + j] * code[is];
}

/* ------------------------------------------------------------------------ */
/* 7. ESS-FWI epoch with the bracketing/parabolic line search.              */
/* ------------------------------------------------------------------------ */

typedef struct {
  orc_geom geom;
  float *vel;               /* padded transformed model (updated in place) */
  const float *wlt;         /* [nt] */
  const float *dobs;        /* [ns][nt][ng] */
  float tmin, tmax;         /* UpdateVelOp bounds */
  int nita;                 /* max_iter_select_alpha3 */
  float maxdv;
  orc_rng codes_rng;        /* essRandomCodes(ESS_SEED=1) */
  float *g0, *dir;          /* previous gradient, update direction */
  float preserved_alpha;    /* UpdateSteplenOp::preservedAlpha */
  int preserved_init;
  float updateobj, initobj;
  /* scratch visible to tests: last epoch's raw data */
  float *last_grad;         /* scaled+masked gradient of the last epoch */
  float last_obj1, last_steplen, last_alpha2, last_alpha3, last_obj2, last_obj3;
  int n_trial_runs;         /* number of calobjval forward runs in the last epoch */
  int last_parabolic;
  /* line-search context */
  const float *ls_encsrc, *ls_encobs;
} orc_ess;

size_t orc_ess_sizeof(void) { return sizeof(orc_ess); }

void orc_ess_init(orc_ess *e) {
  size_t n = (size_t)e->geom.nx * e->geom.nz;
  e->g0 = (float *)calloc(n, sizeof(float));
  e->dir = (float *)calloc(n, sizeof(float));
  e->last_grad = (float *)calloc(n, sizeof(float));
  e->preserved_alpha = 0;
  e->preserved_init = 0;
  e->updateobj = e->initobj = 0;
  orc_rng_seed(&e->codes_rng, 1);
}

void orc_ess_free(orc_ess *e) { free(e->g0); free(e->dir); free(e->last_grad); }

/* Scripted objective for the line-search tests: when set, calobjval returns
 * hook(steplen, ctx) instead of running a trial forward model, so that every
 * branch of refineAlpha can be driven by a hand-made objective table. */
typedef float (*orc_objective_hook)(float steplen, void *ctx);
static orc_objective_hook orc_obj_hook = 0;
static void *orc_obj_hook_ctx = 0;
void orc_set_objective_hook(orc_objective_hook f, void *ctx) { orc_obj_hook = f; orc_obj_hook_ctx = ctx; }

/* UpdateSteplenOp::calobjval (updatesteplenop.cpp:59-86): the mute uses the
 * trial model; shot < 0 = ESS (mute by source 0), else FWI shot. */
static float orc_calobjval(orc_ess *e, const float *dir, float steplen, const float *encsrc, const float *encobs,
                           int shot) {
  if (orc_obj_hook) {
    e->n_trial_runs++;
    return orc_obj_hook(steplen, orc_obj_hook_ctx);
  }
  const orc_geom *g = &e->geom;
  size_t n = (size_t)g->nx * g->nz, m = (size_t)g->nt * g->ng;
  float *nv = (float *)malloc(sizeof(float) * n);
  orc_update_vel(nv, e->vel, dir, n, steplen, e->tmin, e->tmax);
  float *dcal = (float *)calloc(m, sizeof(float));
  orc_forward_modeling(g, nv, encsrc, shot, dcal);
  orc_remove_direct_arrival(g, nv, shot < 0 ? 0 : shot, dcal);
  for (size_t j = 0; j < m; j++) dcal[j] = encobs[j] - dcal[j];
  float val = orc_cal_objective(dcal, m);
  free(nv); free(dcal);
  e->n_trial_runs++;
  return val;
}

/* std::min(a, b) as libstdc++ defines it: (b < a) ? b : a. */
static inline float orc_minf(float a, float b) { return (b < a) ? b : a; }

/* parabolicLessComp-ordered set (updatesteplenop.cpp:19-21,107): only its
 * begin() is ever read, and begin() moves to a newly inserted point iff
 * new.obj - begin.obj < 1e-10 (float difference vs double constant). */
typedef struct { float a, o; int any; } orc_best;
static void orc_best_insert(orc_best *b, float a, float o) {
  if (!b->any || (o - b->o) < 1e-10) { b->a = a; b->o = o; b->any = 1; }
}

/* refineAlpha (updatesteplenop.cpp:88-214), including the two fallbacks that
 * evaluate the local alpha3/alpha2 rather than the returned ones. */
static int orc_refine_alpha(orc_ess *e, float obj_val1, float maxAlpha3, float *_alpha2, float *_obj_val2,
                            float *_alpha3, float *_obj_val3) {
  const float *es = e->ls_encsrc, *eo = e->ls_encobs;
  float alpha3 = *_alpha3, alpha2 = *_alpha2;
  float obj_val2, obj_val3 = 0;
  obj_val2 = orc_calobjval(e, e->dir, alpha2, es, eo, -1);
  obj_val3 = orc_calobjval(e, e->dir, alpha3, es, eo, -1);
  orc_best tuned = {0, 0, 0};
  orc_best_insert(&tuned, alpha2, obj_val2);
  int iter = 0;
  for (; iter < e->nita && obj_val2 > obj_val1; iter++) {
    alpha3 = alpha2;
    obj_val3 = obj_val2;
    alpha2 /= 2;
    obj_val2 = orc_calobjval(e, e->dir, alpha2, es, eo, -1);
    orc_best_insert(&tuned, alpha2, obj_val2);
  }
  if (obj_val2 > obj_val1) {
    *_alpha2 = tuned.a;
    *_obj_val2 = tuned.o;
    *_alpha3 = orc_minf(*_alpha2 * 2, maxAlpha3);
    *_obj_val3 = orc_calobjval(e, e->dir, alpha3, es, eo, -1);
    return 0;
  }
  const float alpha1 = 0;
  float linearFitAlph3 = (obj_val2 - obj_val1) / (alpha2 - alpha1) * (alpha3 - alpha1) + obj_val1;
  tuned.any = 0;
  orc_best_insert(&tuned, alpha3, obj_val3);
  while (obj_val3 < linearFitAlph3 && obj_val3 < obj_val1 && alpha3 < maxAlpha3) {
    alpha2 = alpha3;
    obj_val2 = obj_val3;
    alpha3 = orc_minf(alpha3 * 2, maxAlpha3);
    obj_val3 = orc_calobjval(e, e->dir, alpha3, es, eo, -1);
    orc_best_insert(&tuned, alpha3, obj_val3);
  }
  if (alpha3 > maxAlpha3 + 0.1) {
    *_alpha3 = tuned.a;
    *_obj_val3 = tuned.o;
    *_alpha2 = *_alpha3 / 2;
    *_obj_val2 = orc_calobjval(e, e->dir, alpha2, es, eo, -1);
    return 0;
  }
  *_alpha2 = alpha2; *_alpha3 = alpha3; *_obj_val2 = obj_val2; *_obj_val3 = obj_val3;
  return 1;
}

/* initAlpha23 (updatesteplenop.cpp:273-285) */
static void orc_init_alpha23(orc_ess *e, float maxAlpha3, float *a2, float *a3) {
  const float minAlpha = 1.0E-7, resetAlpha = 1.0E-4;
  if (!e->preserved_init) { e->preserved_init = 1; e->preserved_alpha = maxAlpha3; }
  *a3 = e->preserved_alpha;
  *a3 = *a3 < minAlpha ? resetAlpha : *a3;
  *a2 = *a3 * 0.5;
}

/* calsteplen after calMaxAlpha2_3 (updatesteplenop.cpp:231-264) */
static void orc_steplen_search(orc_ess *e, float obj_val1, float max_alpha3, float *steplen, float *objval) {
  float alpha1 = 0, alpha2, alpha3;
  orc_init_alpha23(e, max_alpha3, &alpha2, &alpha3);
  float obj_val2, obj_val3;
  int toParabolic = orc_refine_alpha(e, obj_val1, max_alpha3, &alpha2, &obj_val2, &alpha3, &obj_val3);
  float alpha4, obj_val4;
  if (toParabolic) {
    orc_parabola_vertex(alpha1, obj_val1, alpha2, obj_val2, alpha3, obj_val3, max_alpha3, &alpha4, &obj_val4);
    if (alpha4 > max_alpha3) alpha4 = max_alpha3;
  } else {
    alpha4 = alpha3;
    obj_val4 = obj_val3;
  }
  e->preserved_alpha = alpha4;
  *steplen = alpha4;
  *objval = obj_val4;
  e->last_alpha2 = alpha2; e->last_alpha3 = alpha3; e->last_obj2 = obj_val2; e->last_obj3 = obj_val3;
  e->last_parabolic = toParabolic;
}

/* calsteplen (updatesteplenop.cpp:216-265) */
static void orc_calsteplen(orc_ess *e, float obj_val1, float *steplen, float *objval) {
  const orc_geom *g = &e->geom;
  size_t n = (size_t)g->nx * g->nz;
  float max_alpha2, max_alpha3;
  orc_cal_max_alpha23(e->vel, e->dir, n, g->dt, g->dx, e->maxdv, &max_alpha2, &max_alpha3);
  orc_steplen_search(e, obj_val1, max_alpha3, steplen, objval);
}

/* One step-length search on a scripted objective (orc_set_objective_hook),
 * for the line-search tests: state[0] = preservedAlpha.alpha, state[1] =
 * preservedAlpha.init on entry, both updated on return; out[10] as
 * orc_ess_scalars (obj1, steplen, alpha2, alpha3, obj2, obj3, objval, -,
 * trial evaluations, parabolic). */
void orc_ess_scalars(const orc_ess *e, float *out);
void orc_steplen_scripted(float obj_val1, float max_alpha3, int nita, float *state, float *out) {
  orc_ess e;
  memset(&e, 0, sizeof e);
  e.nita = nita;
  e.preserved_alpha = state[0];
  e.preserved_init = state[1] != 0;
  float steplen, objval;
  orc_steplen_search(&e, obj_val1, max_alpha3, &steplen, &objval);
  state[0] = e.preserved_alpha;
  state[1] = (float)e.preserved_init;
  e.last_obj1 = obj_val1; e.last_steplen = steplen; e.updateobj = objval;
  orc_ess_scalars(&e, out);
}

/* EssFwiFramework::epoch (essfwiframework.cpp:191-239) */
void orc_ess_epoch(orc_ess *e, int iter) {
  const orc_geom *g = &e->geom;
  int ns = g->ns, nt = g->nt, ng = g->ng;
  size_t n = (size_t)g->nx * g->nz, m = (size_t)nt * ng;
  int *codes = (int *)malloc(sizeof(int) * ns);
  orc_gen_codes(&e->codes_rng, ns, codes);
  float *encsrc = (float *)malloc(sizeof(float) * (size_t)nt * ns);
  float *encobs = (float *)malloc(sizeof(float) * m);
  orc_encode_source(codes, ns, e->wlt, nt, encsrc);
  orc_encode_obs(codes, ns, e->dobs, nt, ng, encobs);
  float *dcal = (float *)calloc(m, sizeof(float));
  orc_forward_modeling(g, e->vel, encsrc, -1, dcal);
  orc_remove_direct_arrival(g, e->vel, 0, encobs);
  orc_remove_direct_arrival(g, e->vel, 0, dcal);
  float *vsrc = (float *)malloc(sizeof(float) * m);
  for (size_t j = 0; j < m; j++) vsrc[j] = encobs[j] - dcal[j];
  float obj1 = orc_cal_objective(vsrc, m);
  e->initobj = iter == 0 ? obj1 : e->initobj;
  orc_trans_vsrc(vsrc, nt, ng);
  float *g1 = (float *)calloc(n, sizeof(float));
  orc_calgradient(g, e->vel, encsrc, -1, vsrc, g1);
  orc_scale_gradient(g, g1);
  orc_mask_gradient(g, g1);
  memcpy(e->last_grad, g1, sizeof(float) * n);
  orc_update_grad(e->g0, g1, e->dir, n, iter);
  e->ls_encsrc = encsrc;
  e->ls_encobs = encobs;
  e->n_trial_runs = 0;
  float steplen;
  orc_calsteplen(e, obj1, &steplen, &e->updateobj);
  orc_update_vel(e->vel, e->vel, e->dir, n, steplen, e->tmin, e->tmax);
  orc_refill_boundary(g, e->vel);
  e->last_obj1 = obj1;
  e->last_steplen = steplen;
  free(codes); free(encsrc); free(encobs); free(dcal); free(vsrc); free(g1);
}

/* ------------------------------------------------------------------------ */
/* 8. Conventional FWI epoch (fwiframework.cpp:269-356).                    */
/* ------------------------------------------------------------------------ */

/* The gradient half of FwiFramework::epoch is followed exactly: per shot,
 * FwiForwardModeling, fwiRemoveDirectArrival on both data, residual, objective
 * (accumulated over shots), transVsrc, calgradient, scale, mask, sum.
 *
 * The step-length half of the reference is undefined behaviour
 * (fwiupdatesteplenop.cpp:88-217 never writes obj_val2/3 back; epoch sums
 * uninitialised members; calobjval reads wlt with stride ns).  The defined
 * rule used here and in the product: one bracket (alpha2, alpha3) from
 * calMaxAlpha2_3/initAlpha23 on the summed direction, obj2 and obj3 as sums
 * over shots of per-shot objectives of the trial model (each shot's own
 * source, both data muted with that shot's window from the trial model for
 * dcal and the current model for dobs), then parabolaVertex clamped to
 * max_alpha3 as FwiUpdateSteplenOp::parabola_fit does. */
void orc_fwi_epoch(orc_ess *e, int iter) {
  const orc_geom *g = &e->geom;
  int ns = g->ns, nt = g->nt, ng = g->ng;
  size_t n = (size_t)g->nx * g->nz, m = (size_t)nt * ng;
  float *g2 = (float *)calloc(n, sizeof(float));
  float *obs = (float *)malloc(sizeof(float) * m), *dcal = (float *)malloc(sizeof(float) * m);
  float *g1 = (float *)malloc(sizeof(float) * n);
  float obj1 = 0.0f;
  for (int is = 0; is < ns; is++) {
    memcpy(obs, &e->dobs[(size_t)is * m], sizeof(float) * m);
    memset(dcal, 0, sizeof(float) * m);
    orc_forward_modeling(g, e->vel, e->wlt, is, dcal);
    orc_remove_direct_arrival(g, e->vel, is, obs);
    orc_remove_direct_arrival(g, e->vel, is, dcal);
    for (size_t j = 0; j < m; j++) dcal[j] = obs[j] - dcal[j];
    obj1 += orc_cal_objective(dcal, m);
    orc_trans_vsrc(dcal, nt, ng);
    memset(g1, 0, sizeof(float) * n);
    orc_calgradient(g, e->vel, e->wlt, is, dcal, g1);
    orc_scale_gradient(g, g1);
    orc_mask_gradient(g, g1);
    for (size_t i = 0; i < n; i++) g2[i] = g2[i] + g1[i];
  }
  e->initobj = iter == 0 ? obj1 : e->initobj;
  memcpy(e->last_grad, g2, sizeof(float) * n);
  orc_update_grad(e->g0, g2, e->dir, n, iter);

  float max_alpha2, max_alpha3, alpha2, alpha3;
  orc_cal_max_alpha23(e->vel, e->dir, n, g->dt, g->dx, e->maxdv, &max_alpha2, &max_alpha3);
  orc_init_alpha23(e, max_alpha3, &alpha2, &alpha3);
  float obj2 = 0, obj3 = 0;
  e->n_trial_runs = 0;
  for (int pass = 0; pass < 2; pass++) {
    float a = pass == 0 ? alpha2 : alpha3;
    float *nv = (float *)malloc(sizeof(float) * n);
    orc_update_vel(nv, e->vel, e->dir, n, a, e->tmin, e->tmax);
    float tot = 0;
    for (int is = 0; is < ns; is++) {
      memcpy(obs, &e->dobs[(size_t)is * m], sizeof(float) * m);
      memset(dcal, 0, sizeof(float) * m);
      orc_forward_modeling(g, nv, e->wlt, is, dcal);
      orc_remove_direct_arrival(g, e->vel, is, obs);
      orc_remove_direct_arrival(g, nv, is, dcal);
      for (size_t j = 0; j < m; j++) dcal[j] = obs[j] - dcal[j];
      tot += orc_cal_objective(dcal, m);
      e->n_trial_runs++;
    }
    if (pass == 0) obj2 = tot; else obj3 = tot;
    free(nv);
  }
  float alpha4, obj4;
  orc_parabola_vertex(0, obj1, alpha2, obj2, alpha3, obj3, max_alpha3, &alpha4, &obj4);
  if (alpha4 > max_alpha3) alpha4 = max_alpha3;
  e->preserved_alpha = alpha4;
  e->updateobj = obj4;
  orc_update_vel(e->vel, e->vel, e->dir, n, alpha4, e->tmin, e->tmax);
  orc_refill_boundary(g, e->vel);
  e->last_obj1 = obj1; e->last_steplen = alpha4;
  e->last_alpha2 = alpha2; e->last_alpha3 = alpha3; e->last_obj2 = obj2; e->last_obj3 = obj3;
  e->last_parabolic = 1;
  free(g2); free(obs); free(dcal); free(g1);
}

/* ------------------------------------------------------------------------ */
/* 9. EnKF pieces that are not linear algebra.                              */
/* ------------------------------------------------------------------------ */

/* clipPosition (Matrix.cpp:170-185) over singular values s[0..n) in FP64. */
int orc_clip_position(const double *s, int n) {
  const double relativeError = 0.001f;
  double sum = 0;
  int pos;
  for (pos = 0; pos < n - 1; pos++) {
    sum += fabs(s[pos]);
    if (relativeError / (1 - relativeError) * sum > (n - pos + 1) * s[pos + 1]) break;
  }
  return pos + 1;
}

/* pCreateAMean (enkfanalyze.cpp:578-597) on one process: FP32 sum over
 * members then divide. */
void orc_create_amean(const float *const *vels, int N, long n, float *mean) {
  for (long i = 0; i < n; i++) {
    float s = 0;
    for (int j = 0; j < N; j++) s += vels[j][i];
    mean[i] = s / N;
  }
}

/* ------------------------------------------------------------------------ */
/* 10. Handle helpers for the ctypes binding used by the tests.             */
/* ------------------------------------------------------------------------ */

/* A geometry with padded dims from physical (nx, nz, nb), as the
 * Damp4t10d ctor + expandDomain produce them (damp4t10d.cpp:138-151,367-376);
 * positions are physical cell coordinates and are copied. */
orc_geom *orc_geom_new(int nx, int nz, int nb, float dt, float dx, float fm, int nt, int ns, const int *sx,
                       const int *sz, int ng, const int *gx, const int *gz, int parallel_u2) {
  orc_geom *g = (orc_geom *)calloc(1, sizeof(orc_geom));
  g->nx = nx + 2 * nb + 2 * EXFD;
  g->nz = nz + nb + 2 * EXFD;
  g->bz0 = EXFD;
  g->bx0 = g->bxn = g->bzn = nb + EXFD;
  g->dt = dt; g->dx = dx; g->fm = fm; g->nt = nt;
  int *p = (int *)malloc(sizeof(int) * (2 * ns + 2 * ng + 4));
  memcpy(p, sx, sizeof(int) * ns); memcpy(p + ns, sz, sizeof(int) * ns);
  memcpy(p + 2 * ns, gx, sizeof(int) * ng); memcpy(p + 2 * ns + ng, gz, sizeof(int) * ng);
  g->ns = ns; g->sx = p; g->sz = p + ns;
  g->ng = ng; g->gx = p + 2 * ns; g->gz = p + 2 * ns + ng;
  g->parallel_u2 = parallel_u2;
  return g;
}
void orc_geom_free(orc_geom *g) { free((void *)g->sx); free(g); }
void orc_geom_dims(const orc_geom *g, int *out) {
  out[0] = g->nx; out[1] = g->nz; out[2] = g->bx0; out[3] = g->bz0; out[4] = orc_bndr_size(g);
}
void orc_geom_set_parallel_u2(orc_geom *g, int p) { g->parallel_u2 = p; }

void orc_write_bndr(const orc_geom *g, float *bndr, const float *p) { orc_rw_bndr(g, bndr, (float *)p, 1); }
void orc_read_bndr(const orc_geom *g, const float *bndr, float *p) { orc_rw_bndr(g, (float *)bndr, p, 0); }
void orc_add_sources(const orc_geom *g, float *p, const float *src, int lo, int hi, int sign) {
  orc_manip(g, p, src, g->sx, g->sz, lo, hi, sign);
}
void orc_add_receivers(const orc_geom *g, float *p, const float *src, int sign) {
  orc_manip(g, p, src, g->gx, g->gz, 0, g->ng, sign);
}
void orc_record_seis(const orc_geom *g, float *seis, const float *p) { orc_record(g, seis, p); }

orc_ess *orc_ess_new(const orc_geom *g, float *vel, const float *wlt, const float *dobs, float vmin, float vmax,
                     int nita, float maxdv) {
  orc_ess *e = (orc_ess *)calloc(1, sizeof(orc_ess));
  e->geom = *g;
  e->vel = vel; e->wlt = wlt; e->dobs = dobs;
  orc_vel_bounds(vmin, vmax, g->dx, g->dt, &e->tmin, &e->tmax);
  e->nita = nita; e->maxdv = maxdv;
  orc_ess_init(e);
  return e;
}
void orc_ess_delete(orc_ess *e) { orc_ess_free(e); free(e); }
/* obj1, steplen, alpha2, alpha3, obj2, obj3, updateobj, initobj, ntrial, parabolic */
void orc_ess_scalars(const orc_ess *e, float *out) {
  out[0] = e->last_obj1; out[1] = e->last_steplen; out[2] = e->last_alpha2; out[3] = e->last_alpha3;
  out[4] = e->last_obj2; out[5] = e->last_obj3; out[6] = e->updateobj; out[7] = e->initobj;
  out[8] = (float)e->n_trial_runs; out[9] = (float)e->last_parabolic;
}
void orc_ess_copy_grad(const orc_ess *e, float *out) {
  memcpy(out, e->last_grad, sizeof(float) * (size_t)e->geom.nx * e->geom.nz);
}
void orc_ess_copy_dir(const orc_ess *e, float *out) {
  memcpy(out, e->dir, sizeof(float) * (size_t)e->geom.nx * e->geom.nz);
}
