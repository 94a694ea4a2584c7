/* essfwiframework.h -- EssFwiFramework (essfwiframework.h:16-53 of swfwi):
 * encoded-source FWI epochs, run device-resident; the host Velocity bound to
 * the Damp4t10d is refreshed after every epoch as the reference updates it
 * in place. */
#ifndef SWFWI_ESSFWIFRAMEWORK_H_
#define SWFWI_ESSFWIFRAMEWORK_H_

#include <cstdio>
#include <memory>
#include <vector>

#include "damp4t10d.h"
#include "updatesteplenop.h"
#include "updatevelop.h"

namespace swfwi {
class Inversion;
}

class EssFwiFramework {
 public:
  EssFwiFramework(Damp4t10d &fmMethod, const UpdateSteplenOp &updateSteplenOp, const UpdateVelOp &updateVelOp,
                  const std::vector<float> &wlt, const std::vector<float> &dobs);
  ~EssFwiFramework();
  void epoch(int iter);
  void writeVel(FILE *file) const;
  float getUpdateObj() const;
  float getInitObj() const;

 protected:
  Damp4t10d &fmMethod;
  std::unique_ptr<swfwi::Inversion> inv;
  bool ess;
};

/* FwiFramework (fwiframework.h:16-53): per-shot gradients summed, all shots
 * batched in one launch per step; the step length follows the defined
 * two-point parabolic rule (the reference's is undefined behaviour). */
class FwiFramework : public EssFwiFramework {
 public:
  FwiFramework(Damp4t10d &fmMethod, const UpdateSteplenOp &updateSteplenOp, const UpdateVelOp &updateVelOp,
               const std::vector<float> &wlt, const std::vector<float> &dobs);
};

#endif
