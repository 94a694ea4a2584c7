// Ensemble-Kalman FWI on one GPU (main-enfwi-damp.cpp:406-631): the N
// members' EssFwiFramework epochs run as ONE batched epoch -- every member's
// encoded forward run, gradient and line-search trials share each launch --
// with the EnKF analysis every `niterenkf` iterations and the objective of
// the ensemble-mean model.
#pragma once
#include <vector>

#include "enkf.h"
#include "linesearch.h"
#include "solver.h"

namespace swfwi {

class EnsembleFwi {
 public:
  // Members live in `enkf` (its device member array is updated in place);
  // observed data and wavelet are the ones `enkf` was built with.
  EnsembleFwi(Solver& s, Enkf& enkf, const UpdateParams& p);

  // essfwis[i]->epoch(iter) for every member i (main-enfwi-damp.cpp:563-567)
  void epoch(int iter);
  // calobj of the ensemble-mean model (main-enfwi-damp.cpp:262-283,586-603):
  // pCreateAMean, then an encoded forward run with the first codes of seed
  // 1, mute, and the objective.  mean_host (optional) receives the mean.
  float mean_objective(float* mean_host);

  int members() const { return N_; }
  const std::vector<EpochStats>& stats() const { return st_; }
  int trial_rounds() const { return rounds_; }
  int trials() const { return ntrial_; }
  double grad_ms() const { return grad_ms_; }
  double linesearch_ms() const { return ls_ms_; }
  void download_grads(float* host);  // [N][npts] scaled + masked gradients of the last epoch

 private:
  // objectives of trial models vel[member[k]] + step[k] * dir[member[k]]
  std::vector<float> trial_objectives(const std::vector<int>& member, const std::vector<float>& step);
  void update_grads(int iter);

  Solver& s_;
  Enkf& enkf_;
  int N_;
  UpdateParams p_;
  float tmin_ = 0, tmax_ = 0;
  CodeStream codes_{1};  // ESS_SEED: every member's generator draws the same codes
  std::vector<PreservedAlpha> preserved_;
  std::vector<EpochStats> st_;
  DBuf<float> rv_, g0_, dir_, g1_, encsrc_, encobs_, obs_, dcal_, band_, res_, vsrc_;
  DBuf<float> tvel_, trv_, tobs_, tdcal_, tres_, scratch_;
  DBuf<double> partials_, dsum_;
  Fields w_;
  int rounds_ = 0, ntrial_ = 0;
  double grad_ms_ = 0, ls_ms_ = 0;
};

}  // namespace swfwi
