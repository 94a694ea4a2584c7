/*
 * oracle64.c -- FP64 run of the acoustic propagation and adjoint gradient
 * (what the low-precision checkpoint path is measured against).
 *
 * TEST INFRASTRUCTURE ONLY.  Built into oracle/liboracle.so beside oracle.c;
 * loaded by tests/ and by bench.py's CPU legs, never by the product library.
 *
 * The same discretisation as the reference -- the Zhang Jinhai coefficients
 * exactly as the reference stores them (float, fd4t10s-damp-zjh.c:24-29),
 * the damped leapfrog of fd4t10s-damp-zjh.c:14-115, the undamped one of
 * fd4t10s-zjh.c:10-56, Damp4t10d's injection / recording / boundary band
 * (damp4t10d.cpp:193-214,163-180,447-536) and the calgradient sequence of
 * fwiframework.cpp:114-177 -- evaluated entirely in double precision, fields
 * included.  Its gradient is the reference gradient without FP32 rounding:
 * the yardstick for gradients computed from reduced-precision wavefield
 * checkpoints.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EX64 6

static const float ZJH64[6] = {+1.53400796f, +1.78858721f, -0.31660756f,
                               +0.07612173f, -0.01626042f, +0.00216736f};

typedef struct {
  int nx, nz;             /* padded dims */
  int bx0, bz0, bxn, bzn; /* damp4t10d.cpp:372-373 */
  int nt;
  float dt;
} o64_geom;

/* u2 over [5, n-5) and the damped (nb > 0) or undamped update over
 * [6, n-6) in double (fd4t10s-damp-zjh.c:41-88, fd4t10s-zjh.c:28-53). */
static void o64_step(double *prev, const double *cur, const double *rv, double *u2, int nx, int nz, int nb) {
  const double a0 = ZJH64[0], a1 = ZJH64[1], a2 = ZJH64[2], a3 = ZJH64[3], a4 = ZJH64[4], a5 = ZJH64[5];
#pragma omp parallel for schedule(static)
  for (int ix = EX64 - 1; ix < nx - (EX64 - 1); ix++)
    for (int iz = EX64 - 1; iz < nz - (EX64 - 1); iz++) {
      const long i = (long)ix * nz + iz;
      const double *c = cur;
      u2[i] = -4.0 * a0 * c[i] + a1 * (c[i - 1] + c[i + 1] + c[i - nz] + c[i + nz]) +
              a2 * (c[i - 2] + c[i + 2] + c[i - 2 * nz] + c[i + 2 * nz]) +
              a3 * (c[i - 3] + c[i + 3] + c[i - 3 * nz] + c[i + 3 * nz]) +
              a4 * (c[i - 4] + c[i + 4] + c[i - 4 * nz] + c[i + 4 * nz]) +
              a5 * (c[i - 5] + c[i + 5] + c[i - 5 * nz] + c[i + 5 * nz]);
    }
#pragma omp parallel for schedule(static)
  for (int ix = EX64; ix < nx - EX64; ix++)
    for (int iz = EX64; iz < nz - EX64; iz++) {
      const long i = (long)ix * nz + iz;
      double delta = 0;
      if (nb > 0) { /* fd4t10s-damp-zjh.c:64-80: the z branch last */
        double dist = 0;
        if (ix < nb) dist = (double)(nb - ix) / nb;
        if (ix >= nx - nb) dist = (double)(ix - (nx - nb) + 1) / nb;
        if (iz >= nz - nb) dist = (double)(iz - (nz - nb) + 1) / nb;
        delta = 0.05 * dist * dist;
      }
      const double r = rv[i];
      const double lap = u2[i - 1] + u2[i + 1] + u2[i - nz] + u2[i + nz] - 4.0 * u2[i];
      prev[i] = (2.0 - 2.0 * delta + delta * delta) * cur[i] - (1.0 - 2.0 * delta) * prev[i] + r * u2[i] +
                r * r / 12.0 * lap;
    }
}

static int o64_bndr_size(const o64_geom *g) {
  const int w = 6;
  int nxb = g->nx - (g->bx0 - w + g->bxn - w);
  int nzb = g->nz - g->bz0 - g->bzn;
  return w * (nxb + 2 * nzb);
}

/* writeBndry / readBndry (damp4t10d.cpp:478-536) */
static void o64_rw_bndr(const o64_geom *g, double *bndr, double *p, int write) {
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

static void o64_geom_init(o64_geom *g, int nxpad, int nzpad, int nb, int nt, float dt) {
  g->nx = nxpad; g->nz = nzpad;
  g->bz0 = EX64; g->bx0 = g->bxn = g->bzn = nb + EX64;
  g->nt = nt; g->dt = dt;
}

/* 1 / vel_t in double: the Courant number squared of each point */
static double *o64_rv(const float *vel, long n) {
  double *rv = (double *)malloc(sizeof(double) * n);
  for (long i = 0; i < n; i++) rv[i] = 1.0 / (double)vel[i];
  return rv;
}

/* FwiForwardModeling of one source (damp4t10d.cpp:272-289): inject
 * src[it] at (sx, sz), step, swap, record the receivers gx[r], gz[r]
 * (physical coordinates) into dcal[it*ng + r].  vel is the padded
 * transformed model of nxpad x nzpad floats. */
void o64_forward_modeling(int nxpad, int nzpad, int nb, int nt, float dt, const float *vel, const double *src,
                          int sx, int sz, int ng, const int *gx, const int *gz, double *dcal) {
  o64_geom g;
  o64_geom_init(&g, nxpad, nzpad, nb, nt, dt);
  const long n = (long)nxpad * nzpad;
  double *rv = o64_rv(vel, n);
  double *p0 = (double *)calloc(n, sizeof(double)), *p1 = (double *)calloc(n, sizeof(double));
  double *u2 = (double *)calloc(n, sizeof(double));
  const long si = (long)(sx + g.bx0) * nzpad + (sz + g.bz0);
  for (int it = 0; it < nt; it++) {
    p1[si] += src[it];
    o64_step(p0, p1, rv, u2, nxpad, nzpad, g.bx0);
    double *t = p0; p0 = p1; p1 = t;
    for (int r = 0; r < ng; r++) dcal[(long)it * ng + r] = p0[(long)(gx[r] + g.bx0) * nzpad + (gz[r] + g.bz0)];
  }
  free(rv); free(p0); free(p1); free(u2);
}

/* calgradient of one source (fwiframework.cpp:114-177) in double: forward
 * with writeBndry, backward with readBndry, undamped reconstruction,
 * subSource(src[it]), receiver injection of vsrc[it*ng + r], damped receiver
 * step, and g -= s * r * w(it) with the reference's float/double taper test.
 * g (padded grid) is accumulated into. */
void o64_calgradient(int nxpad, int nzpad, int nb, int nt, float dt, const float *vel, const double *src, int sx,
                     int sz, int ng, const int *gx, const int *gz, const double *vsrc, double *grad) {
  o64_geom g;
  o64_geom_init(&g, nxpad, nzpad, nb, nt, dt);
  const long n = (long)nxpad * nzpad;
  const int bs = o64_bndr_size(&g);
  double *rv = o64_rv(vel, n);
  double *bndr = (double *)calloc((size_t)nt * bs, sizeof(double));
  double *sp0 = (double *)calloc(n, sizeof(double)), *sp1 = (double *)calloc(n, sizeof(double));
  double *gp0 = (double *)calloc(n, sizeof(double)), *gp1 = (double *)calloc(n, sizeof(double));
  double *u2 = (double *)calloc(n, sizeof(double));
  const long si = (long)(sx + g.bx0) * nzpad + (sz + g.bz0);
  for (int it = 0; it < nt; it++) {
    sp1[si] += src[it];
    o64_step(sp0, sp1, rv, u2, nxpad, nzpad, g.bx0);
    double *t = sp0; sp0 = sp1; sp1 = t;
    o64_rw_bndr(&g, &bndr[(size_t)it * bs], sp0, 1);
  }
  for (int it = nt - 1; it >= 0; it--) {
    o64_rw_bndr(&g, &bndr[(size_t)it * bs], sp0, 0);
    double *t = sp0; sp0 = sp1; sp1 = t;
    o64_step(sp0, sp1, rv, u2, nxpad, nzpad, 0);
    sp0[si] -= src[it];
    for (int r = 0; r < ng; r++) gp1[(long)(gx[r] + g.bx0) * nzpad + (gz[r] + g.bz0)] += vsrc[(long)it * ng + r];
    o64_step(gp0, gp1, rv, u2, nxpad, nzpad, g.bx0);
    t = gp0; gp0 = gp1; gp1 = t;
    double w;
    if (dt * it > 0.4) w = 1.0;
    else if (dt * it > 0.3) w = (dt * it - 0.3) / 0.1;
    else break;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) grad[i] -= sp0[i] * gp0[i] * w;
  }
  free(rv); free(bndr); free(sp0); free(sp1); free(gp0); free(gp1); free(u2);
}
