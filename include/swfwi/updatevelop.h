/* updatevelop.h -- UpdateVelOp (updatevelop.h:13-21 of swfwi): clamped model
 * update in transformed units. */
#ifndef SWFWI_UPDATEVELOP_H_
#define SWFWI_UPDATEVELOP_H_

#include <vector>

#include "velocity.h"

class UpdateVelOp {
 public:
  UpdateVelOp(float vmin, float vmax, float dx, float dt);
  void update(Velocity &newVel, const Velocity &vel, const std::vector<float> &grad, float steplen) const;
  /* the m/s bounds given to the constructor (for the device drivers) */
  float vminPhysical() const { return vmin_ms; }
  float vmaxPhysical() const { return vmax_ms; }

 private:
  float vmin;
  float vmax;
  float vmin_ms, vmax_ms;
};

typedef UpdateVelOp FwiUpdateVelOp; /* fwiupdatevelop.h:13-21: same operation */

#endif
