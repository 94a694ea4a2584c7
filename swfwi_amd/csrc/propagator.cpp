#include "propagator.h"

#include <algorithm>
#include <numeric>
#include <string>

namespace swfwi {

void Geometry::derive() {
  if (nx <= 0 || nz <= 0 || nt <= 0) fail(kErrArg, "grid and time dimensions must be positive");
  if (nb < 6) fail(kErrArg, "sponge width nb must be >= 6 (the 6-wide boundary band lives inside the sponge)");
  if (!(dt > 0) || !(dx > 0)) fail(kErrArg, "dt and dx must be positive");
  if (sx.size() != sz.size() || gx.size() != gz.size()) fail(kErrArg, "position arrays differ in length");
  if (sx.empty() || gx.empty()) fail(kErrArg, "need at least one source and one receiver");
  // Params::check (main-fm-damp.cpp:183-194): every position inside the model
  for (size_t i = 0; i < sx.size(); i++)
    if (sx[i] < 0 || sz[i] < 0 || sx[i] >= nx || sz[i] >= nz) fail(kErrArg, "sources exceeds the computing zone");
  for (size_t i = 0; i < gx.size(); i++)
    if (gx[i] < 0 || gz[i] < 0 || gx[i] >= nx || gz[i] >= nz) fail(kErrArg, "geophones exceeds the computing zone");
  nxpad = nx + 2 * nb + 2 * kHalo;
  nzpad = nz + nb + 2 * kHalo;
  ld = nzpad;
  bz0 = kHalo;
  bx0 = bxn = bzn = nb + kHalo;
  const int w = 6;
  const int nxb = nxpad - (bx0 - w + bxn - w);
  const int nzb = nzpad - bz0 - bzn;
  bsize = w * (nxb + 2 * nzb);
  ntz = (nzpad - 2 * kHalo + kTZ - 1) / kTZ;
  ntx = (nxpad - 2 * kHalo + kTX - 1) / kTX;
}

void build_damping(int nx, int nz, int nb, DampTables& t, hipStream_t s) {
  // fd4t10s-damp-zjh.c:64-80 with bx = bz = nb; the bottom branch wins.
  const float max_delta = 0.05f;
  std::vector<float> dxv(nx), dzv(nz, 0.f);
  for (int ix = 0; ix < nx; ix++) {
    float dist = 0;
    if (ix < nb) dist = (float)(nb - ix) / nb;
    if (ix >= nx - nb) dist = (float)(ix - (nx - nb) + 1) / nb;
    dxv[ix] = max_delta * dist * dist;
  }
  for (int iz = 0; iz < nz; iz++) {
    if (iz >= nz - nb) {
      float dist = (float)(iz - (nz - nb) + 1) / nb;
      dzv[iz] = max_delta * dist * dist;
    }
  }
  t.dx.upload(dxv, s);
  t.dz.upload(dzv, s);
  t.zdamp = nz - nb;
}

void build_points(const std::vector<long long>& idx, int ld, int ntz, int ntx, DevPoints& out, hipStream_t s) {
  const int ntiles = ntz * ntx;
  const int n = (int)idx.size();
  std::vector<int> tile(n);
  for (int p = 0; p < n; p++) {
    const int ix = (int)(idx[p] / ld), iz = (int)(idx[p] % ld);
    const int tz = (iz - kHalo) / kTZ, tx = (ix - kHalo) / kTX;
    if (iz < kHalo || ix < kHalo || tz >= ntz || tx >= ntx) fail(kErrArg, "point outside the stencil update region");
    tile[p] = tx * ntz + tz;
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return tile[a] < tile[b]; });
  std::vector<int> ptr(ntiles + 1, 0);
  for (int p = 0; p < n; p++) ptr[tile[p] + 1]++;
  for (int t = 0; t < ntiles; t++) ptr[t + 1] += ptr[t];
  std::vector<int> idx32(n);
  for (int p = 0; p < n; p++) idx32[p] = (int)idx[p];
  out.tile_ptr.upload(ptr, s);
  out.items.upload(order, s);
  out.idx.upload(idx32, s);
  out.n = n;
}

void Fields::ensure(int e, long long npts, bool receiver) {
  E = e;
  const size_t n = (size_t)e * (size_t)npts;
  a.resize(n);
  b.resize(n);
  if (receiver) {
    c.resize(n);
    d.resize(n);
  }
}

Propagator::Propagator(const Geometry& g, int precision) : g_(g), prec_(precision) {
  g_.derive();
  if (precision != kExact && precision != kFp32) fail(kErrArg, "unknown precision mode");
  hipStream_t s = stream_.get();
  build_damping(g_.nxpad, g_.nzpad, g_.bx0, damp_, s);
  std::vector<long long> si(g_.ns()), gi(g_.ng());
  std::vector<int> sxp(g_.ns()), szp(g_.ns()), gxp(g_.ng()), gzp(g_.ng());
  for (int i = 0; i < g_.ns(); i++) {
    sxp[i] = g_.sx[i] + g_.bx0;
    szp[i] = g_.sz[i] + g_.bz0;
    si[i] = (long long)sxp[i] * g_.ld + szp[i];
  }
  for (int i = 0; i < g_.ng(); i++) {
    gxp[i] = g_.gx[i] + g_.bx0;
    gzp[i] = g_.gz[i] + g_.bz0;
    gi[i] = (long long)gxp[i] * g_.ld + gzp[i];
  }
  build_points(si, g_.ld, g_.ntz, g_.ntx, src_, s);
  build_points(gi, g_.ld, g_.ntz, g_.ntx, rec_, s);
  sxp_.upload(sxp, s);
  szp_.upload(szp, s);
  gxp_.upload(gxp, s);
  gzp_.upload(gzp, s);
  stream_.sync();
}

Grid Propagator::grid() const {
  Grid gr;
  gr.nx = g_.nxpad;
  gr.nz = g_.nzpad;
  gr.ld = g_.ld;
  gr.ntz = g_.ntz;
  gr.ntx = g_.ntx;
  gr.nb = g_.bx0;
  gr.bx0 = g_.bx0;
  gr.bxn = g_.bxn;
  gr.bz0 = g_.bz0;
  gr.bzn = g_.bzn;
  gr.delta_x = damp_.dx.get();
  gr.delta_z = damp_.dz.get();
  gr.zdamp = damp_.zdamp;
  return gr;
}

int Propagator::xcorr_range(float dt, int nt, std::vector<float>& weights) {
  // fwiframework.cpp:164-175: float dt*it against double thresholds; the loop
  // breaks at the first it without an imaging condition.
  weights.assign(nt, 0.f);
  int it_min = nt;
  for (int it = nt - 1; it >= 0; it--) {
    if (dt * it > 0.4) {
      weights[it] = 1.0;
    } else if (dt * it > 0.3) {
      weights[it] = (dt * it - 0.3) / 0.1;
    } else {
      break;
    }
    it_min = it;
  }
  return it_min;
}

void Propagator::forward(int E, const float* rv, long long rvstride, const Sources& src, float* dcal, float* band,
                         Fields& w, int rvdiv) {
  hipStream_t s = stream();
  const long long np = g_.npts();
  const int nt = g_.nt, ng = g_.ng();
  w.ensure(E, np, false);
  launch_fill(w.a.get(), (long long)E * np, 0.f, s);
  launch_fill(w.b.get(), (long long)E * np, 0.f, s);

  Inject inj;
  inj.pts = src_.view();
  inj.amp = src.amp;
  inj.ld = src.ld;
  inj.estride = src.estride;
  inj.sign = +1;
  inj.lo = src.lo;
  inj.hi = src.hi;

  // step 0 injects into the initially zero cur (addSource before stepForward)
  inj.t = 0;
  launch_point_inject(w.b.get(), np, inj, g_.ns(), E, s);

  FwdStepArgs a;
  a.g = grid();
  a.fstride = np;
  a.rv = rv;
  a.rvstride = rvstride;
  a.rvdiv = rvdiv;
  a.inj = inj;
  a.rec.pts = rec_.view();
  a.rec.estride = (long long)nt * ng;
  a.band_estride = (long long)nt * g_.bsize;
  for (int it = 0; it < nt; it++) {
    float* cur = (it % 2 == 0) ? w.b.get() : w.a.get();
    float* prv = (it % 2 == 0) ? w.a.get() : w.b.get();
    a.cur = cur;
    a.prev = prv;
    a.inj.t = it + 1 < nt ? it + 1 : -1;
    a.rec.out = dcal ? dcal + (long long)it * ng : nullptr;
    a.band = band ? band + (long long)it * g_.bsize : nullptr;
    launch_fwd_step(a, E, true, prec_, s);
    fwd_launches++;
  }
  w.s_cur = ((nt - 1) % 2 == 0) ? w.b.get() : w.a.get();
  w.s_nxt = ((nt - 1) % 2 == 0) ? w.a.get() : w.b.get();
  SWFWI_HIP(hipGetLastError());
}

int Propagator::backward(int E, const float* rv, long long rvstride, const Sources& src, const float* vsrc,
                         long long vstride, const float* band, float* grad, long long gstride, Fields& w,
                         int rvdiv) {
  hipStream_t s = stream();
  const long long np = g_.npts();
  const int nt = g_.nt, ng = g_.ng();
  if (!w.s_cur) fail(kErrState, "backward() needs the state of a forward() run");
  w.ensure(E, np, true);
  std::vector<float> wts;
  const int it_min = xcorr_range(g_.dt, nt, wts);
  if (it_min >= nt) return 0;

  float* s_cur = w.s_cur;
  float* s_prv = w.s_nxt;
  float* g_cur = w.c.get();
  float* g_prv = w.d.get();
  launch_fill(g_cur, (long long)E * np, 0.f, s);
  launch_fill(g_prv, (long long)E * np, 0.f, s);

  Inject ginj;
  ginj.pts = rec_.view();
  ginj.amp = vsrc;
  ginj.ld = ng;
  ginj.estride = vstride;
  ginj.sign = +1;
  ginj.t = nt - 1;
  launch_point_inject(g_cur, np, ginj, ng, E, s);

  BwdStepArgs a;
  a.g = grid();
  a.fstride = np;
  a.rv = rv;
  a.rvstride = rvstride;
  a.rvdiv = rvdiv;
  a.s_sub.pts = src_.view();
  a.s_sub.amp = src.amp;
  a.s_sub.ld = src.ld;
  a.s_sub.estride = src.estride;
  a.s_sub.sign = -1;
  a.s_sub.lo = src.lo;
  a.s_sub.hi = src.hi;
  a.g_inj = ginj;
  a.band_estride = (long long)nt * g_.bsize;
  a.grad = grad;
  a.gstride = gstride;
  a.do_xcorr = 1;
  int iters = 0;
  for (int it = nt - 1; it >= it_min; it--) {
    a.s_cur = s_cur;
    a.s_prev = s_prv;
    a.g_cur = g_cur;
    a.g_prev = g_prv;
    a.s_sub.t = it;
    const bool more = it - 1 >= it_min;
    a.g_inj.t = more ? it - 1 : -1;
    a.band = more ? band + (long long)(it - 1) * g_.bsize : nullptr;
    a.weight = wts[it];
    launch_bwd_step(a, E, prec_, s);
    bwd_launches++;
    std::swap(s_cur, s_prv);
    std::swap(g_cur, g_prv);
    iters++;
  }
  SWFWI_HIP(hipGetLastError());
  return iters;
}

}  // namespace swfwi
