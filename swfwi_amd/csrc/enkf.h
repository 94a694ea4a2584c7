// Ensemble Kalman analysis on the GPU (EnkfAnalyze, enkfanalyze.h:21-58):
// batched member forward runs, FP64 data-space products, an FP64 N x N
// eigen-solve on the host, and the fused per-point model update.
#pragma once
#include <cstdint>
#include <vector>

#include "solver.h"

namespace swfwi {

constexpr int kMaxMembers = 1024;

// device pieces (enkf.hip)
void enkf_ha_perturb(const float* ha, int N, long long nd, double* hap, double* max_abs_value, hipStream_t s);
void enkf_band(const float* ha, const double* hap, const float* d, const double* e_in, uint64_t seed, double sigma,
               int N, long long nd, double* band, double* t0, hipStream_t s);
void enkf_model_update(const float* vels, float* out, long long n, int N, const double* T, float dx, float dt,
                       hipStream_t s);
void enkf_member_mean(const float* vels, long long n, int N, float* mean, hipStream_t s);

// Symmetric eigen-decomposition (cyclic Jacobi, FP64): a is N x N column
// major (destroyed); eigenvalues descending in w, eigenvectors in the columns
// of v.  Returns false if the sweeps did not converge.
bool jacobi_eigen(std::vector<double>& a, int n, std::vector<double>& w, std::vector<double>& v);
// clipPosition (Matrix.cpp:170-185) over descending singular values.
int clip_position(const std::vector<double>& s);

class Enkf {
 public:
  Enkf(Solver& s, const float* wlt_host, const float* dobs_host, float sigfac, int nmembers);

  int members() const { return N_; }
  void set_members(const float* host);  // [N][nxpad][nzpad], transformed
  void get_members(float* host);
  float* members_dev() { return mem_.get(); }
  const float* dobs_dev() const { return dobs_.get(); }  // [ns][nt][ng]
  // EnkfAnalyze::pAnalyze: perturbation injected ([N][nt][ng], FP64) or, when
  // null, drawn N(0, sigma) from the counter-based generator with `seed`.
  void analyze(const double* e_host, uint64_t seed);
  // pCreateAMean: FP32 member-order mean model
  void mean(float* host);

  double sigma() const { return sigma_; }
  int clip() const { return clip_; }
  const std::vector<double>& svals() const { return svals_; }
  const std::vector<double>& gain() const { return t4_; }
  CodeStream& codes() { return codes_; }

 private:
  Solver& s_;
  int N_;
  float sigfac_;
  DBuf<float> dobs_, mem_, mem_out_, rv_, ha_, d_, encsrc_, mean_;
  DBuf<int> code_;
  DBuf<double> hap_, band_, t0_, e_, G_, P_, Q_, T_;
  CodeStream codes_{2};  // ENKF_SEED (enkfanalyze.h:44)
  bool sigma_init_ = false;
  double sigma_ = 0;
  int clip_ = 0;
  std::vector<double> svals_, t4_;
  Fields w_;
};

}  // namespace swfwi
