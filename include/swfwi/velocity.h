/* velocity.h -- Velocity (velocity.h:13-21 of swfwi): a model on an nx x nz
 * grid, index ix*nz + iz. */
#ifndef SWFWI_VELOCITY_H_
#define SWFWI_VELOCITY_H_

#include <vector>

class Velocity {
 public:
  Velocity(int _nx, int _nz) : dat((size_t)_nx * _nz, 0), nx(_nx), nz(_nz) {}
  Velocity(const std::vector<float> &_dat, int _nx, int _nz) : dat(_dat), nx(_nx), nz(_nz) {}

 public:
  std::vector<float> dat;
  int nx;
  int nz;
};

#endif
