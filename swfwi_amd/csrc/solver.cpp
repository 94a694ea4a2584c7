#include "solver.h"

#include "linesearch.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>

namespace swfwi {

namespace {

// Upload a small host vector and wait for it (the vector is a temporary).
template <typename T>
void upload_now(DBuf<T>& d, const std::vector<T>& h, hipStream_t s) {
  d.upload(h, s);
  SWFWI_HIP(hipStreamSynchronize(s));
}

template <typename T>
T read_scalar(const T* dptr, hipStream_t s) {
  T v;
  SWFWI_HIP(hipMemcpyAsync(&v, dptr, sizeof(T), hipMemcpyDeviceToHost, s));
  SWFWI_HIP(hipStreamSynchronize(s));
  return v;
}

}  // namespace

Solver::Solver(const Geometry& g, int precision) : prop_(g, precision) {}

void Solver::bind_velocity(const float* host_vel) {
  vel_.upload(host_vel, (size_t)npts(), stream());
  rv_.resize((size_t)npts());
  refresh_rv();
  prop_.sync();
  bound_ = true;
}

void Solver::download_velocity(float* host_vel) {
  if (!bound_) fail(kErrState, "bind the velocity first");
  vel_.download(host_vel, (size_t)npts(), stream());
  prop_.sync();
}

void Solver::refresh_rv() { op_reciprocal(vel_.get(), rv_.get(), npts(), stream()); }

void Solver::set_wavelet(const float* wlt_host) {
  const Geometry& g = geo();
  std::vector<float> w(wlt_host, wlt_host + g.nt);
  std::vector<float> tab((size_t)g.nt * g.ns());
  for (int t = 0; t < g.nt; t++)
    for (int s = 0; s < g.ns(); s++) tab[(size_t)t * g.ns() + s] = w[t];
  wlt_.upload(w, stream());
  wtab_.upload(tab, stream());
  prop_.sync();
}

double Solver::point_updates() const { return prop_.fwd_point_updates + prop_.bwd_point_updates; }

void Solver::forward_modeling(const float* encsrc, int shot, float* dcal_host) {
  if (!bound_) fail(kErrState, "bind the velocity first");
  const Geometry& g = geo();
  hipStream_t s = stream();
  const int ns = g.ns(), nt = g.nt, ng = g.ng();
  if (shot >= ns) fail(kErrArg, "shot id out of range");
  std::vector<float> tab((size_t)nt * ns, 0.f);
  if (shot >= 0) {
    for (int t = 0; t < nt; t++) tab[(size_t)t * ns + shot] = encsrc[t];
    upload_now(lo, std::vector<int>{shot}, s);
    upload_now(hi, std::vector<int>{shot + 1}, s);
  } else {
    std::memcpy(tab.data(), encsrc, sizeof(float) * tab.size());
  }
  upload_now(this->encsrc, tab, s);
  Propagator::Sources src{this->encsrc.get(), ns, 0, shot >= 0 ? lo.get() : nullptr, shot >= 0 ? hi.get() : nullptr};
  dcal.resize((size_t)nt * ng);
  prop_.forward(1, rv_.get(), 0, src, dcal.get(), nullptr, w);
  dcal.download(dcal_host, (size_t)nt * ng, s);
  prop_.sync();
}

namespace {
// 256-byte aligned float count of a workspace region
size_t ws_round(size_t n) { return (n + 63) / 64 * 64; }
}  // namespace

size_t Solver::workspace_floats() const {
  const Geometry& g = geo();
  // two fields, the recorded seismogram, the source table, lo/hi ids
  return 2 * ws_round((size_t)npts()) + ws_round((size_t)g.nt * g.ng()) + ws_round((size_t)g.nt * g.ns()) +
         ws_round(2);
}

void Solver::forward_modeling_ws(const float* encsrc_host, int shot, float* dcal_host, float* ws, size_t ws_floats) {
  if (!bound_) fail(kErrState, "bind the velocity first");
  const Geometry& g = geo();
  const int ns = g.ns(), nt = g.nt, ng = g.ng();
  if (shot >= ns) fail(kErrArg, "shot id out of range");
  const size_t need = workspace_floats();
  if (need > ws_floats)
    fail(kErrRange, "forward modelling needs " + std::to_string(need * sizeof(float)) +
                        " bytes of workspace, more than the " + std::to_string(ws_floats * sizeof(float)) +
                        " available");
  hipStream_t s = stream();
  float* fa = ws;
  float* fb = fa + ws_round((size_t)npts());
  float* dc = fb + ws_round((size_t)npts());
  float* tab = dc + ws_round((size_t)nt * ng);
  int* lohi = reinterpret_cast<int*>(tab + ws_round((size_t)nt * ns));
  const int ids[2] = {shot, shot + 1};
  if (shot >= 0) {
    // encsrc[t] into column `shot` of the [nt][ns] table; the other columns
    // are never read (the entry injects source ids [shot, shot+1) only)
    SWFWI_HIP(hipMemcpy2DAsync(tab + shot, sizeof(float) * ns, encsrc_host, sizeof(float), sizeof(float), nt,
                               hipMemcpyHostToDevice, s));
    SWFWI_HIP(hipMemcpyAsync(lohi, ids, sizeof(ids), hipMemcpyHostToDevice, s));
  } else {
    SWFWI_HIP(hipMemcpyAsync(tab, encsrc_host, sizeof(float) * nt * ns, hipMemcpyHostToDevice, s));
  }
  Propagator::Sources src{tab, ns, 0, shot >= 0 ? lohi : nullptr, shot >= 0 ? lohi + 1 : nullptr};
  prop_.forward_external(rv_.get(), src, dc, fa, fb);
  SWFWI_HIP(hipMemcpyAsync(dcal_host, dc, sizeof(float) * nt * ng, hipMemcpyDeviceToHost, s));
  // the host arrays (ids included) must outlive the copies
  prop_.sync();
}

void Solver::calgradient(const float* encsrc, int shot, const float* vsrc_host, float* grad_host) {
  if (!bound_) fail(kErrState, "bind the velocity first");
  const Geometry& g = geo();
  hipStream_t s = stream();
  const int ns = g.ns(), nt = g.nt, ng = g.ng();
  if (shot >= ns) fail(kErrArg, "shot id out of range");
  std::vector<float> tab((size_t)nt * ns, 0.f);
  if (shot >= 0) {
    for (int t = 0; t < nt; t++) tab[(size_t)t * ns + shot] = encsrc[t];
    upload_now(lo, std::vector<int>{shot}, s);
    upload_now(hi, std::vector<int>{shot + 1}, s);
  } else {
    std::memcpy(tab.data(), encsrc, sizeof(float) * tab.size());
  }
  upload_now(this->encsrc, tab, s);
  Propagator::Sources src{this->encsrc.get(), ns, 0, shot >= 0 ? lo.get() : nullptr, shot >= 0 ? hi.get() : nullptr};
  vsrc.upload(vsrc_host, (size_t)nt * ng, s);
  band.resize((size_t)nt * g.bsize);
  grad.resize((size_t)npts());
  launch_fill(grad.get(), npts(), 0.f, s);
  prop_.forward(1, rv_.get(), 0, src, nullptr, band.get(), w);
  prop_.backward(1, rv_.get(), 0, src, vsrc.get(), 0, band.get(), grad.get(), 0, w);
  grad.download(grad_host, (size_t)npts(), s);
  prop_.sync();
}

void Solver::forward_shots(const float* wlt_host, float* dobs_host) {
  // fm-damp's per-shot loop (main-fm-damp.cpp:241-256): every shot in one
  // batch, entry e injecting source e with amplitude wlt[it].
  if (!bound_) fail(kErrState, "bind the velocity first");
  const Geometry& g = geo();
  hipStream_t s = stream();
  const int ns = g.ns();
  const long long m = (long long)g.nt * g.ng();
  set_wavelet(wlt_host);
  std::vector<int> l(ns), h(ns);
  for (int e = 0; e < ns; e++) {
    l[e] = e;
    h[e] = e + 1;
  }
  upload_now(lo, l, s);
  upload_now(hi, h, s);
  Propagator::Sources src{wtab_.get(), ns, 0, lo.get(), hi.get()};
  dcal.resize((size_t)ns * m);
  prop_.forward(ns, rv_.get(), 0, src, dcal.get(), nullptr, w);
  dcal.download(dobs_host, (size_t)ns * m, s);
  prop_.sync();
}

void Solver::mute(const float* vel, long long vstride, const std::vector<int>& srcs, float* data, long long dstride) {
  const Geometry& g = geo();
  hipStream_t s = stream();
  const int E = (int)srcs.size();
  std::vector<int> sx(E), sz(E);
  for (int e = 0; e < E; e++) {
    sx[e] = g.sx[srcs[e]] + g.bx0;
    sz[e] = g.sz[srcs[e]] + g.bz0;
  }
  std::vector<int> both(sx);
  both.insert(both.end(), sz.begin(), sz.end());
  upload_now(srcid, both, s);
  const float t_width = 1.5 / g.fm;  // damp4t10d.cpp:390-399
  const int half_len = t_width / g.dt;
  win_start.resize((size_t)E * g.ng());
  win_end.resize((size_t)E * g.ng());
  op_mute_windows(vel, vstride, E, g.gridgeom(), srcid.get(), srcid.get() + E, prop_.rec_xp(), prop_.rec_zp(), g.ng(),
                  g.nt, half_len, win_start.get(), win_end.get(), s);
  op_mute_apply(data, dstride, E, win_start.get(), win_end.get(), g.ng(), g.nt, g.ng(), s);
}

std::vector<float> Solver::residual(const float* obs, long long ostride, const float* cal, float* resbuf, int E) {
  const Geometry& g = geo();
  hipStream_t s = stream();
  const long long m = (long long)g.nt * g.ng();
  partials.resize((size_t)E * 512);
  dsum.resize((size_t)E);
  const int nb = op_residual(obs, ostride, cal, resbuf, m, E, partials.get(), s);
  op_reduce_partials(partials.get(), nb, E, dsum.get(), s);
  std::vector<double> h(E);
  dsum.download(h.data(), E, s);
  SWFWI_HIP(hipStreamSynchronize(s));
  std::vector<float> out(E);
  for (int e = 0; e < E; e++) out[e] = (float)h[e];
  return out;
}

// ---------------------------------------------------------------------------

Inversion::Inversion(Solver& s, const float* wlt_host, const float* dobs_host, const UpdateParams& p) : s_(s), p_(p) {
  const Geometry& g = s.geo();
  // UpdateVelOp ctor (updatevelop.cpp:11-21 of the listing: lines 11-21)
  tmin_ = (g.dx / g.dt / p.vmax) * (g.dx / g.dt / p.vmax);
  tmax_ = (g.dx / g.dt / p.vmin) * (g.dx / g.dt / p.vmin);
  if (tmax_ <= tmin_) fail(kErrArg, "vmax must exceed vmin");
  // epochs re-run the same time loops on the same buffers: replay graphs
  s.prop().set_graphs(std::max(s.prop().graphs_mode(), 1));
  s.set_wavelet(wlt_host);
  const size_t m = (size_t)g.ns() * g.nt * g.ng();
  dobs_.upload(dobs_host, m, s.stream());
  g0_.resize((size_t)s.npts());
  dir_.resize((size_t)s.npts());
  g1_.resize((size_t)s.npts());
  launch_fill(g0_.get(), s.npts(), 0.f, s.stream());
  launch_fill(dir_.get(), s.npts(), 0.f, s.stream());
  s.prop().sync();
}

void Inversion::download_grad(float* out) {
  g1_.download(out, (size_t)s_.npts(), s_.stream());
  s_.prop().sync();
}
void Inversion::download_direction(float* out) {
  dir_.download(out, (size_t)s_.npts(), s_.stream());
  s_.prop().sync();
}

void Inversion::update_grad(const float* g1, int iter) {
  // updateGrad, PR+ (fwiframework.cpp:42-72), dots accumulated in FP64
  hipStream_t s = s_.stream();
  const long long n = s_.npts();
  if (iter == 0) {
    SWFWI_HIP(hipMemcpyAsync(dir_.get(), g1, n * sizeof(float), hipMemcpyDeviceToDevice, s));
    SWFWI_HIP(hipMemcpyAsync(g0_.get(), g1, n * sizeof(float), hipMemcpyDeviceToDevice, s));
    return;
  }
  s_.partials.resize(3 * 512);
  s_.dsum.resize(3);
  const int nb = op_cg_dots(g1, g0_.get(), n, s_.partials.get(), s);
  op_reduce_partials(s_.partials.get(), nb, 3, s_.dsum.get(), s);
  double h[3];
  s_.dsum.download(h, 3, s);
  SWFWI_HIP(hipStreamSynchronize(s));
  const float a = (float)h[0], b = (float)h[1], c = (float)h[2];
  float beta = (a - b) / c;
  if (beta < 0.0f) beta = 0.0f;
  op_cg_update(g1, g0_.get(), dir_.get(), beta, n, s);
}

float Inversion::max_alpha3() {
  const Geometry& g = s_.geo();
  s_.fscratch.resize(1024);
  op_max_alpha(s_.vel(), dir_.get(), s_.npts(), g.dt, g.dx, p_.maxdv, s_.fscratch.get(), s_.fscratch.get() + 1023,
               s_.stream());
  const float a2 = read_scalar(s_.fscratch.get() + 1023, s_.stream());
  return 2 * a2;
}

std::vector<float> Inversion::calobjvals(const std::vector<float>& steps) {
  // UpdateSteplenOp::calobjval (updatesteplenop.cpp:59-86) for several trial
  // step lengths in one batched launch sequence.
  const Geometry& g = s_.geo();
  hipStream_t s = s_.stream();
  const int E = (int)steps.size();
  const long long n = s_.npts(), m = (long long)g.nt * g.ng();
  s_.trial_vel.resize((size_t)E * n);
  s_.trial_rv.resize((size_t)E * n);
  op_update_vel(s_.vel(), dir_.get(), steps.data(), E, n, tmin_, tmax_, s_.trial_vel.get(), n, s);
  op_reciprocal(s_.trial_vel.get(), s_.trial_rv.get(), (long long)E * n, s);
  s_.dcal.resize((size_t)E * m);
  Propagator::Sources src{s_.encsrc.get(), g.ns(), 0, nullptr, nullptr};
  s_.prop().forward(E, s_.trial_rv.get(), n, src, s_.dcal.get(), nullptr, s_.w);
  s_.mute(s_.trial_vel.get(), n, std::vector<int>(E, 0), s_.dcal.get(), m);
  s_.res.resize((size_t)E * m);
  std::vector<float> o = s_.residual(s_.encobs.get(), 0, s_.dcal.get(), s_.res.get(), E);
  st_.ntrial += E;
  return o;
}

void Inversion::apply_update(float steplen) {
  // UpdateVelOp::update in place + refillBoundary (essfwiframework.cpp:235-238)
  const Geometry& g = s_.geo();
  hipStream_t s = s_.stream();
  op_update_vel(s_.vel(), dir_.get(), &steplen, 1, s_.npts(), tmin_, tmax_, s_.vel(), 0, s);
  op_refill(s_.vel(), 0, 1, g.gridgeom(), s);
  s_.refresh_rv();
}

void Inversion::ess_epoch(int iter) {
  const Geometry& g = s_.geo();
  hipStream_t s = s_.stream();
  const int ns = g.ns(), nt = g.nt, ng = g.ng();
  const long long n = s_.npts(), m = (long long)nt * ng;
  hipEvent_t e0, e1, e2;
  SWFWI_HIP(hipEventCreate(&e0));
  SWFWI_HIP(hipEventCreate(&e1));
  SWFWI_HIP(hipEventCreate(&e2));
  SWFWI_HIP(hipEventRecord(e0, s));
  st_.ntrial = 0;

  // new +-1 codes every epoch (essfwiframework.cpp:193-201)
  const std::vector<int> code = codes_.next(ns);
  upload_now(s_.codes, code, s);
  s_.encsrc.resize((size_t)nt * ns);
  s_.encobs.resize((size_t)m);
  op_encode_source(s_.wavelet(), s_.codes.get(), ns, nt, s_.encsrc.get(), s);
  op_encode_obs(dobs_.get(), s_.codes.get(), ns, m, s_.encobs.get(), s);

  // EssForwardModeling and calgradient's forward half are the same run on the
  // same model: one pass records dcal and saves the boundary band.
  s_.dcal.resize((size_t)m);
  s_.band.resize((size_t)nt * g.bsize);
  Propagator::Sources src{s_.encsrc.get(), ns, 0, nullptr, nullptr};
  s_.prop().forward(1, s_.rv(), 0, src, s_.dcal.get(), s_.band.get(), s_.w);
  s_.mute(s_.vel(), 0, {0}, s_.encobs.get(), 0);
  s_.mute(s_.vel(), 0, {0}, s_.dcal.get(), 0);
  s_.res.resize((size_t)m);
  const float obj1 = s_.residual(s_.encobs.get(), 0, s_.dcal.get(), s_.res.get(), 1)[0];
  st_.initobj = iter == 0 ? obj1 : st_.initobj;
  s_.vsrc.resize((size_t)m);
  op_trans_vsrc(s_.res.get(), s_.vsrc.get(), nt, ng, 1, s);
  launch_fill(g1_.get(), n, 0.f, s);
  s_.prop().backward(1, s_.rv(), 0, src, s_.vsrc.get(), 0, s_.band.get(), g1_.get(), 0, s_.w);
  op_scale_mask(g1_.get(), 0, 1, g.gridgeom(), s);
  update_grad(g1_.get(), iter);
  SWFWI_HIP(hipEventRecord(e1, s));

  // UpdateSteplenOp::calsteplen (updatesteplenop.cpp:216-265); the first two
  // trials run as one batch of two models
  const float max3 = max_alpha3();
  float alpha2, alpha3;
  preserved_.init_alpha23(max3, alpha2, alpha3);
  RefineAlpha ra(obj1, max3, alpha2, alpha3, p_.nita);
  while (!ra.done()) {
    const std::vector<float> o = calobjvals(ra.pending());
    ra.feed(o.data());
  }
  float alpha4, obj_val4;
  final_step(ra, obj1, max3, alpha4, obj_val4);
  const int toParabolic = ra.parabolic;
  alpha2 = ra.ra2;
  alpha3 = ra.ra3;
  const float obj_val2 = ra.ro2, obj_val3 = ra.ro3;
  preserved_.alpha = alpha4;
  apply_update(alpha4);
  SWFWI_HIP(hipEventRecord(e2, s));
  SWFWI_HIP(hipEventSynchronize(e2));
  float ms1 = 0, ms2 = 0;
  SWFWI_HIP(hipEventElapsedTime(&ms1, e0, e1));
  SWFWI_HIP(hipEventElapsedTime(&ms2, e1, e2));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  st_.obj1 = obj1;
  st_.steplen = alpha4;
  st_.alpha2 = alpha2;
  st_.alpha3 = alpha3;
  st_.obj2 = obj_val2;
  st_.obj3 = obj_val3;
  st_.updateobj = obj_val4;
  st_.parabolic = toParabolic;
  st_.grad_ms = ms1;
  st_.linesearch_ms = ms2;
}

void Inversion::fwi_epoch(int iter) {
  const Geometry& g = s_.geo();
  hipStream_t s = s_.stream();
  const int ns = g.ns(), nt = g.nt, ng = g.ng();
  const long long n = s_.npts(), m = (long long)nt * ng;
  hipEvent_t e0, e1, e2;
  SWFWI_HIP(hipEventCreate(&e0));
  SWFWI_HIP(hipEventCreate(&e1));
  SWFWI_HIP(hipEventCreate(&e2));
  SWFWI_HIP(hipEventRecord(e0, s));
  st_.ntrial = 0;

  // every shot in one batch: entry e = shot e, injecting source e alone
  std::vector<int> lo(2 * ns), hi(2 * ns), own(ns);
  for (int e = 0; e < 2 * ns; e++) {
    lo[e] = e % ns;
    hi[e] = e % ns + 1;
  }
  for (int e = 0; e < ns; e++) own[e] = e;
  upload_now(s_.lo, lo, s);
  upload_now(s_.hi, hi, s);
  Propagator::Sources src{s_.wavelet_table(), ns, 0, s_.lo.get(), s_.hi.get()};

  // observed data of each shot muted with its own window (fwiframework.cpp:276-281)
  s_.tmp.resize((size_t)ns * m);
  SWFWI_HIP(hipMemcpyAsync(s_.tmp.get(), dobs_.get(), (size_t)ns * m * sizeof(float), hipMemcpyDeviceToDevice, s));
  s_.mute(s_.vel(), 0, own, s_.tmp.get(), m);
  s_.dcal.resize((size_t)2 * ns * m);
  s_.band.resize((size_t)ns * nt * g.bsize);
  s_.prop().forward(ns, s_.rv(), 0, src, s_.dcal.get(), s_.band.get(), s_.w);
  s_.mute(s_.vel(), 0, own, s_.dcal.get(), m);
  s_.res.resize((size_t)2 * ns * m);
  const std::vector<float> objs = s_.residual(s_.tmp.get(), m, s_.dcal.get(), s_.res.get(), ns);
  float obj1 = 0.0f;
  for (int e = 0; e < ns; e++) obj1 += objs[e];
  st_.initobj = iter == 0 ? obj1 : st_.initobj;
  s_.vsrc.resize((size_t)ns * m);
  op_trans_vsrc(s_.res.get(), s_.vsrc.get(), nt, ng, ns, s);
  s_.grad.resize((size_t)ns * n);
  launch_fill(s_.grad.get(), (long long)ns * n, 0.f, s);
  s_.prop().backward(ns, s_.rv(), 0, src, s_.vsrc.get(), m, s_.band.get(), s_.grad.get(), n, s_.w);
  op_scale_mask(s_.grad.get(), n, ns, g.gridgeom(), s);
  op_sum_entries(s_.grad.get(), n, ns, n, g1_.get(), s);
  update_grad(g1_.get(), iter);
  SWFWI_HIP(hipEventRecord(e1, s));

  // defined step rule: one (alpha2, alpha3) bracket, both trial models for
  // every shot in one batch, per-shot objectives summed, parabola
  const float max3 = max_alpha3();
  float alpha2, alpha3;
  preserved_.init_alpha23(max3, alpha2, alpha3);
  const float steps[2] = {alpha2, alpha3};
  s_.trial_vel.resize((size_t)2 * n);
  s_.trial_rv.resize((size_t)2 * n);
  op_update_vel(s_.vel(), dir_.get(), steps, 2, n, tmin_, tmax_, s_.trial_vel.get(), n, s);
  op_reciprocal(s_.trial_vel.get(), s_.trial_rv.get(), 2 * n, s);
  s_.prop().forward(2 * ns, s_.trial_rv.get(), n, src, s_.dcal.get(), nullptr, s_.w, ns);
  std::vector<int> trial_src(2 * ns);
  for (int e = 0; e < 2 * ns; e++) trial_src[e] = e % ns;
  for (int k = 0; k < 2; k++)
    s_.mute(s_.trial_vel.get() + (long long)k * n, 0, own, s_.dcal.get() + (long long)k * ns * m, m);
  float obj[2] = {0, 0};
  for (int k = 0; k < 2; k++) {
    const std::vector<float> o =
        s_.residual(s_.tmp.get(), m, s_.dcal.get() + (long long)k * ns * m, s_.res.get(), ns);
    for (int e = 0; e < ns; e++) obj[k] += o[e];
  }
  st_.ntrial = 2 * ns;
  float alpha4, obj4;
  parabola_vertex(0, obj1, alpha2, obj[0], alpha3, obj[1], alpha4, obj4);
  if (alpha4 > max3) alpha4 = max3;
  preserved_.alpha = alpha4;
  apply_update(alpha4);
  SWFWI_HIP(hipEventRecord(e2, s));
  SWFWI_HIP(hipEventSynchronize(e2));
  float ms1 = 0, ms2 = 0;
  SWFWI_HIP(hipEventElapsedTime(&ms1, e0, e1));
  SWFWI_HIP(hipEventElapsedTime(&ms2, e1, e2));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  st_.obj1 = obj1;
  st_.steplen = alpha4;
  st_.alpha2 = alpha2;
  st_.alpha3 = alpha3;
  st_.obj2 = obj[0];
  st_.obj3 = obj[1];
  st_.updateobj = obj4;
  st_.parabolic = 1;
  st_.grad_ms = ms1;
  st_.linesearch_ms = ms2;
}

}  // namespace swfwi
