/* enkfanalyze.h -- EnkfAnalyze (enkfanalyze.h:21-58 of swfwi) on one GPU. */
#ifndef SWFWI_ENKFANALYZE_H_
#define SWFWI_ENKFANALYZE_H_

#include <memory>
#include <vector>

#include "damp4t10d.h"

namespace swfwi {
class Enkf;
}

class EnkfAnalyze {
 public:
  EnkfAnalyze(const Damp4t10d &fm, const std::vector<float> &wlt, const std::vector<float> &dobs, float sigmafactor);
  ~EnkfAnalyze();
  /* velSet: N pointers to padded transformed models, updated in place */
  void pAnalyze(std::vector<float *> &velSet) const;
  std::vector<float> pCreateAMean(const std::vector<float *> &velSet, const int N) const;
  /* seed of the perturbation generator (the reference seeds boost::mt19937
   * with the MPI rank) */
  void setPerturbationSeed(unsigned long long seed) { seed_ = seed; }

 private:
  const Damp4t10d &fm;
  const std::vector<float> &wlt;
  const std::vector<float> &dobs;
  float sigmaFactor;
  mutable std::unique_ptr<swfwi::Enkf> impl;
  mutable unsigned long long seed_ = 0;
  mutable unsigned long long calls_ = 0;
};

#endif
