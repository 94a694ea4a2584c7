#include "ensemble.h"

#include <algorithm>
#include <cmath>

namespace swfwi {

EnsembleFwi::EnsembleFwi(Solver& s, Enkf& enkf, const UpdateParams& p)
    : s_(s), enkf_(enkf), N_(enkf.members()), p_(p), preserved_(enkf.members()), st_(enkf.members()) {
  const Geometry& g = s.geo();
  // UpdateVelOp ctor (updatevelop.cpp:11-21)
  tmin_ = (g.dx / g.dt / p.vmax) * (g.dx / g.dt / p.vmax);
  tmax_ = (g.dx / g.dt / p.vmin) * (g.dx / g.dt / p.vmin);
  if (tmax_ <= tmin_) fail(kErrArg, "vmax must exceed vmin");
  const size_t n = (size_t)s.npts() * N_;
  g0_.resize(n);
  dir_.resize(n);
  g1_.resize(n);
  launch_fill(g0_.get(), (long long)n, 0.f, s.stream());
  launch_fill(dir_.get(), (long long)n, 0.f, s.stream());
  s.prop().sync();
}

void EnsembleFwi::download_grads(float* host) {
  g1_.download(host, (size_t)N_ * s_.npts(), s_.stream());
  s_.prop().sync();
}

void EnsembleFwi::update_grads(int iter) {
  // updateGrad per member (essfwiframework.cpp:42-72), FP64 dots
  hipStream_t s = s_.stream();
  const long long n = s_.npts();
  if (iter == 0) {
    const size_t bytes = (size_t)N_ * n * sizeof(float);
    SWFWI_HIP(hipMemcpyAsync(dir_.get(), g1_.get(), bytes, hipMemcpyDeviceToDevice, s));
    SWFWI_HIP(hipMemcpyAsync(g0_.get(), g1_.get(), bytes, hipMemcpyDeviceToDevice, s));
    return;
  }
  partials_.resize((size_t)N_ * 3 * 512);
  dsum_.resize((size_t)N_ * 3);
  for (int i = 0; i < N_; i++) {
    const int nb = op_cg_dots(g1_.get() + i * n, g0_.get() + i * n, n, partials_.get() + (size_t)i * 3 * 512, s);
    op_reduce_partials(partials_.get() + (size_t)i * 3 * 512, nb, 3, dsum_.get() + 3 * i, s);
  }
  std::vector<double> h((size_t)N_ * 3);
  dsum_.download(h.data(), h.size(), s);
  SWFWI_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < N_; i++) {
    const float a = (float)h[3 * i], b = (float)h[3 * i + 1], c = (float)h[3 * i + 2];
    float beta = (a - b) / c;
    if (beta < 0.0f) beta = 0.0f;
    op_cg_update(g1_.get() + i * n, g0_.get() + i * n, dir_.get() + i * n, beta, n, s);
  }
}

std::vector<float> EnsembleFwi::trial_objectives(const std::vector<int>& member, const std::vector<float>& step) {
  // UpdateSteplenOp::calobjval (updatesteplenop.cpp:59-86) of every member's
  // pending trial in one batched forward run
  const Geometry& g = s_.geo();
  hipStream_t s = s_.stream();
  const int E = (int)member.size();
  const long long n = s_.npts(), m = (long long)g.nt * g.ng();
  const float* vel = enkf_.members_dev();
  tvel_.resize((size_t)E * n);
  trv_.resize((size_t)E * n);
  tobs_.resize((size_t)E * m);
  for (int k = 0; k < E; k++) {
    const long long off = (long long)member[k] * n;
    op_update_vel(vel + off, dir_.get() + off, &step[k], 1, n, tmin_, tmax_, tvel_.get() + k * n, 0, s);
    SWFWI_HIP(hipMemcpyAsync(tobs_.get() + k * m, obs_.get() + (long long)member[k] * m, m * sizeof(float),
                             hipMemcpyDeviceToDevice, s));
  }
  op_reciprocal(tvel_.get(), trv_.get(), (long long)E * n, s);
  tdcal_.resize((size_t)E * m);
  Propagator::Sources src{encsrc_.get(), g.ns(), 0, nullptr, nullptr};
  s_.prop().forward(E, trv_.get(), n, src, tdcal_.get(), nullptr, w_);
  s_.mute(tvel_.get(), n, std::vector<int>(E, 0), tdcal_.get(), m);
  tres_.resize((size_t)E * m);
  ntrial_ += E;
  rounds_++;
  return s_.residual(tobs_.get(), m, tdcal_.get(), tres_.get(), E);
}

void EnsembleFwi::epoch(int iter) {
  const Geometry& g = s_.geo();
  hipStream_t s = s_.stream();
  const int ns = g.ns(), nt = g.nt, ng = g.ng(), N = N_;
  const long long n = s_.npts(), m = (long long)nt * ng;
  float* vel = enkf_.members_dev();
  hipEvent_t e0, e1, e2;
  SWFWI_HIP(hipEventCreate(&e0));
  SWFWI_HIP(hipEventCreate(&e1));
  SWFWI_HIP(hipEventCreate(&e2));
  SWFWI_HIP(hipEventRecord(e0, s));
  rounds_ = ntrial_ = 0;

  // the epoch's codes, source and observed supershot (essfwiframework.cpp:193-201)
  const std::vector<int> code = codes_.next(ns);
  s_.codes.upload(code, s);
  SWFWI_HIP(hipStreamSynchronize(s));
  encsrc_.resize((size_t)nt * ns);
  encobs_.resize((size_t)m);
  op_encode_source(s_.wavelet(), s_.codes.get(), ns, nt, encsrc_.get(), s);
  op_encode_obs(enkf_.dobs_dev(), s_.codes.get(), ns, m, encobs_.get(), s);

  // every member's forward run with boundary save, one batch
  rv_.resize((size_t)N * n);
  op_reciprocal(vel, rv_.get(), (long long)N * n, s);
  dcal_.resize((size_t)N * m);
  band_.resize((size_t)N * nt * g.bsize);
  Propagator::Sources src{encsrc_.get(), ns, 0, nullptr, nullptr};
  s_.prop().forward(N, rv_.get(), n, src, dcal_.get(), band_.get(), w_);
  // removeDirectArrival with each member's model, on its copy of the
  // observed supershot and on its modelled one
  obs_.resize((size_t)N * m);
  for (int i = 0; i < N; i++)
    SWFWI_HIP(hipMemcpyAsync(obs_.get() + i * m, encobs_.get(), m * sizeof(float), hipMemcpyDeviceToDevice, s));
  const std::vector<int> zero(N, 0);
  s_.mute(vel, n, zero, obs_.get(), m);
  s_.mute(vel, n, zero, dcal_.get(), m);
  res_.resize((size_t)N * m);
  const std::vector<float> obj1 = s_.residual(obs_.get(), m, dcal_.get(), res_.get(), N);
  vsrc_.resize((size_t)N * m);
  op_trans_vsrc(res_.get(), vsrc_.get(), nt, ng, N, s);
  launch_fill(g1_.get(), (long long)N * n, 0.f, s);
  s_.prop().backward(N, rv_.get(), n, src, vsrc_.get(), m, band_.get(), g1_.get(), n, w_);
  op_scale_mask(g1_.get(), n, N, g.gridgeom(), s);
  update_grads(iter);
  SWFWI_HIP(hipEventRecord(e1, s));

  // calsteplen of every member: calMaxAlpha2_3, initAlpha23, then the
  // members' refineAlpha searches advanced in lockstep, each round's
  // pending trials evaluated together
  scratch_.resize((size_t)N * 1024);
  for (int i = 0; i < N; i++)
    op_max_alpha(vel + i * n, dir_.get() + i * n, n, g.dt, g.dx, p_.maxdv, scratch_.get() + (size_t)i * 1024,
                 scratch_.get() + (size_t)i * 1024 + 1023, s);
  std::vector<float> sc((size_t)N * 1024);
  scratch_.download(sc.data(), sc.size(), s);
  SWFWI_HIP(hipStreamSynchronize(s));
  std::vector<float> max3(N);
  std::vector<RefineAlpha> ra(N);
  for (int i = 0; i < N; i++) {
    max3[i] = 2 * sc[(size_t)i * 1024 + 1023];
    float a2, a3;
    preserved_[i].init_alpha23(max3[i], a2, a3);
    ra[i] = RefineAlpha(obj1[i], max3[i], a2, a3, p_.nita);
  }
  std::vector<int> ntrials(N, 0);
  for (;;) {
    std::vector<int> who;
    std::vector<float> steps;
    for (int i = 0; i < N; i++)
      for (float a : ra[i].pending()) {
        who.push_back(i);
        steps.push_back(a);
        ntrials[i]++;
      }
    if (who.empty()) break;
    const std::vector<float> o = trial_objectives(who, steps);
    size_t k = 0;
    for (int i = 0; i < N; i++) {
      const size_t np = ra[i].pending().size();
      if (np) ra[i].feed(o.data() + k);
      k += np;
    }
  }
  // UpdateVelOp::update + refillBoundary per member (essfwiframework.cpp:235-238)
  for (int i = 0; i < N; i++) {
    float alpha4, obj4;
    final_step(ra[i], obj1[i], max3[i], alpha4, obj4);
    preserved_[i].alpha = alpha4;
    op_update_vel(vel + i * n, dir_.get() + i * n, &alpha4, 1, n, tmin_, tmax_, vel + i * n, 0, s);
    EpochStats& st = st_[i];
    st.obj1 = obj1[i];
    st.initobj = iter == 0 ? obj1[i] : st.initobj;
    st.steplen = alpha4;
    st.alpha2 = ra[i].ra2;
    st.alpha3 = ra[i].ra3;
    st.obj2 = ra[i].ro2;
    st.obj3 = ra[i].ro3;
    st.updateobj = obj4;
    st.parabolic = ra[i].parabolic;
    st.ntrial = ntrials[i];
  }
  op_refill(vel, n, N, g.gridgeom(), s);
  SWFWI_HIP(hipEventRecord(e2, s));
  SWFWI_HIP(hipEventSynchronize(e2));
  float ms1 = 0, ms2 = 0;
  SWFWI_HIP(hipEventElapsedTime(&ms1, e0, e1));
  SWFWI_HIP(hipEventElapsedTime(&ms2, e1, e2));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  grad_ms_ = ms1;
  ls_ms_ = ms2;
  for (EpochStats& st : st_) {
    st.grad_ms = ms1;
    st.linesearch_ms = ms2;
  }
}

float EnsembleFwi::mean_objective(float* mean_host) {
  // pCreateAMean, bind, calobj (main-enfwi-damp.cpp:262-283): codes from a
  // fresh seed-1 generator, so always its first draw
  const Geometry& g = s_.geo();
  hipStream_t s = s_.stream();
  const int ns = g.ns(), nt = g.nt, ng = g.ng();
  const long long n = s_.npts(), m = (long long)nt * ng;
  tvel_.resize((size_t)n);
  trv_.resize((size_t)n);
  enkf_member_mean(enkf_.members_dev(), n, N_, tvel_.get(), s);
  op_reciprocal(tvel_.get(), trv_.get(), n, s);
  CodeStream first(1);
  const std::vector<int> code = first.next(ns);
  s_.codes.upload(code, s);
  SWFWI_HIP(hipStreamSynchronize(s));
  encsrc_.resize((size_t)nt * ns);
  tobs_.resize((size_t)m);
  op_encode_source(s_.wavelet(), s_.codes.get(), ns, nt, encsrc_.get(), s);
  op_encode_obs(enkf_.dobs_dev(), s_.codes.get(), ns, m, tobs_.get(), s);
  tdcal_.resize((size_t)m);
  Propagator::Sources src{encsrc_.get(), ns, 0, nullptr, nullptr};
  s_.prop().forward(1, trv_.get(), n, src, tdcal_.get(), nullptr, w_);
  s_.mute(tvel_.get(), 0, {0}, tobs_.get(), 0);
  s_.mute(tvel_.get(), 0, {0}, tdcal_.get(), 0);
  tres_.resize((size_t)m);
  const float obj = s_.residual(tobs_.get(), 0, tdcal_.get(), tres_.get(), 1)[0];
  if (mean_host) {
    tvel_.download(mean_host, (size_t)n, s);
    SWFWI_HIP(hipStreamSynchronize(s));
  }
  return obj;
}

}  // namespace swfwi
