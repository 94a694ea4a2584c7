// Device-resident FWI drivers over the Propagator: forward modelling,
// calgradient, the ESS-FWI and FWI epochs with their line searches, and the
// ensemble (EnKF) forward runs and analysis.  These follow the reference's
// control flow step for step (essfwiframework.cpp:191-239,
// fwiframework.cpp:269-356, updatesteplenop.cpp:59-285,
// enkfanalyze.cpp:113-183,343-495); only scalars cross to the host.
#pragma once
#include <cstdint>
#include <vector>

#include "linesearch.h"
#include "propagator.h"

namespace swfwi {

// boost::minstd_rand + uniform_int<>(0,1) -> +-1 (random-code.cpp:2-22).
struct CodeStream {
  uint64_t x = 1;
  explicit CodeStream(int seed = 1) { reseed(seed); }
  void reseed(int seed) {
    uint64_t s = (uint64_t)(uint32_t)seed % 2147483647u;
    x = s == 0 ? 1 : s;
  }
  std::vector<int> next(int n) {
    std::vector<int> c(n);
    for (int i = 0; i < n; i++) {
      x = (x * 48271u) % 2147483647u;
      c[i] = (int)((x - 1u) / 1073741823u) * 2 - 1;
    }
    return c;
  }
};

// The model plus scratch shared by every driver on one geometry.
class Solver {
 public:
  Solver(const Geometry& g, int precision);

  Propagator& prop() { return prop_; }
  const Geometry& geo() const { return prop_.geo(); }
  hipStream_t stream() const { return prop_.stream(); }
  long long npts() const { return prop_.geo().npts(); }

  // model (padded, transformed units dx^2/(v^2 dt^2))
  void bind_velocity(const float* host_vel);
  void download_velocity(float* host_vel);
  float* vel() { return vel_.get(); }
  const float* rv() const { return rv_.get(); }
  void refresh_rv();

  // Host-array entry points (the handle API): FwiForwardModeling (shot >= 0,
  // amplitude encsrc[it]) / EssForwardModeling (shot < 0, encsrc[it*ns+s]).
  void forward_modeling(const float* encsrc, int shot, float* dcal);
  // calgradient with the same source convention; vsrc [nt][ng]; grad (padded)
  // receives the raw accumulated imaging condition.
  void calgradient(const float* encsrc, int shot, const float* vsrc, float* grad);
  // forward_modeling with every device buffer carved from a caller-provided
  // workspace of `ws_floats` floats (256-byte aligned) instead of the
  // handle's own: allocates nothing.  Fails with kErrRange when the workspace
  // is too small (workspace_floats() gives the need).
  void forward_modeling_ws(const float* encsrc, int shot, float* dcal, float* ws, size_t ws_floats);
  size_t workspace_floats() const;
  // Observed-data generation for every shot (fm-damp): dobs [ns][nt][ng].
  void forward_shots(const float* wlt, float* dobs);

  // Device building blocks used by the drivers.
  // Mute windows of E models (vel + e*vstride) with source `src_of(e)`; apply
  // them to `data` entries.  srcs: host list of source ids per entry.
  void mute(const float* vel, long long vstride, const std::vector<int>& srcs, float* data, long long dstride);
  // objective of E residual entries res[e] = obs(+e*ostride) - cal[e]; returns
  // FP64 sums rounded to float.
  std::vector<float> residual(const float* obs, long long ostride, const float* cal, float* res, int E);
  // sources table for FWI shots: amp[t][ns] = wlt[t] for every column
  void set_wavelet(const float* wlt_host);
  const float* wavelet_table() const { return wtab_.get(); }
  const float* wavelet() const { return wlt_.get(); }

  // throughput accounting
  double point_updates() const;
  long long launches() const { return prop_.fwd_launches + prop_.bwd_launches; }

  // scratch
  Fields w;
  DBuf<float> dcal, band, res, vsrc, grad, trial_vel, trial_rv, tmp;
  DBuf<double> partials, dsum;
  DBuf<int> win_start, win_end, srcid, lo, hi, codes;
  DBuf<float> encsrc, encobs;
  DBuf<float> fscratch;

 private:
  Propagator prop_;
  DBuf<float> vel_, rv_, wlt_, wtab_;
  bool bound_ = false;
};

// Line-search + update parameters (UpdateVelOp, UpdateSteplenOp ctor args).
struct UpdateParams {
  float vmin = 0, vmax = 0;  // m/s (UpdateVelOp ctor, updatevelop.cpp:11-21)
  int nita = 0;              // max_iter_select_alpha3
  float maxdv = 0;
};

struct EpochStats {
  float obj1 = 0, steplen = 0, alpha2 = 0, alpha3 = 0, obj2 = 0, obj3 = 0;
  float updateobj = 0, initobj = 0;
  int ntrial = 0, parabolic = 0;
  double grad_ms = 0, linesearch_ms = 0;
};

// Shared state of UpdateSteplenOp + updateGrad across epochs.
class Inversion {
 public:
  Inversion(Solver& s, const float* wlt_host, const float* dobs_host, const UpdateParams& p);
  // EssFwiFramework::epoch (essfwiframework.cpp:191-239)
  void ess_epoch(int iter);
  // FwiFramework::epoch (fwiframework.cpp:269-356) with the defined step rule
  void fwi_epoch(int iter);
  const EpochStats& stats() const { return st_; }
  float update_obj() const { return st_.updateobj; }
  float init_obj() const { return st_.initobj; }
  void download_grad(float* out);       // scaled + masked gradient of the last epoch
  void download_direction(float* out);
  CodeStream& codes() { return codes_; }

 private:
  // UpdateSteplenOp::calobjval (updatesteplenop.cpp:59-86) for several trial
  // step lengths in one batched forward run
  std::vector<float> calobjvals(const std::vector<float>& steps);
  float max_alpha3();
  void update_grad(const float* g1, int iter);
  void apply_update(float steplen);

  Solver& s_;
  UpdateParams p_;
  float tmin_ = 0, tmax_ = 0;
  CodeStream codes_{1};
  DBuf<float> dobs_, g0_, dir_, g1_;
  PreservedAlpha preserved_;
  EpochStats st_;
};

}  // namespace swfwi
