/* updatesteplenop.h -- UpdateSteplenOp (updatesteplenop.h:15-45 of swfwi).
 * The bracketing/parabolic line search itself runs inside the device
 * drivers (EssFwiFramework / FwiFramework); this class carries its
 * parameters exactly as the reference's constructor takes them. */
#ifndef SWFWI_UPDATESTEPLENOP_H_
#define SWFWI_UPDATESTEPLENOP_H_

#include "damp4t10d.h"
#include "updatevelop.h"

class UpdateSteplenOp {
 public:
  UpdateSteplenOp(const Damp4t10d &fmMethod, const UpdateVelOp &updateVelOp, int max_iter_select_alpha3, float maxdv)
      : fmMethod(&fmMethod), updateVelOp(&updateVelOp), max_iter_select_alpha3(max_iter_select_alpha3), maxdv(maxdv) {}
  int maxIterSelectAlpha3() const { return max_iter_select_alpha3; }
  float maxDv() const { return maxdv; }
  const UpdateVelOp &velOp() const { return *updateVelOp; }
  const Damp4t10d &modeling() const { return *fmMethod; }

 private:
  const Damp4t10d *fmMethod;
  const UpdateVelOp *updateVelOp;
  int max_iter_select_alpha3;
  float maxdv;
};

typedef UpdateSteplenOp FwiUpdateSteplenOp; /* fwiupdatesteplenop.h: same parameters */

#endif
