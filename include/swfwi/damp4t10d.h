/*
 * damp4t10d.h -- drop-in C++ surface of the reference's propagator class
 * (damp4t10d.h:18-87 of swfwi), backed by libswfwi_hip.so.
 *
 * Same class name, constructor and member functions as the reference, so
 * its tool mains and drivers compile against it.  Arrays stay caller-owned
 * host memory as in the reference; the time loops (FwiForwardModeling,
 * EssForwardModeling) run device-resident on the GPU and copy only the
 * seismogram back, the per-step calls (stepForward/stepBackward) run one
 * GPU step each with host copies (compatibility path).
 */
#ifndef SWFWI_DAMP4T10D_H_
#define SWFWI_DAMP4T10D_H_

#include <cstdio>
#include <memory>
#include <vector>

#include "shot-position.h"
#include "velocity.h"

namespace swfwi {
class Solver;
}

class Damp4t10d {
 public:
  Damp4t10d(const ShotPosition &allSrcPos, const ShotPosition &allGeoPos, float dt, float dx, float fm, int nb, int nt);
  Damp4t10d(const Damp4t10d &other);
  Damp4t10d &operator=(const Damp4t10d &other);
  ~Damp4t10d();

  Velocity expandDomain(const Velocity &vel);

  void stepForward(float *p0, float *p1) const;
  void stepBackward(float *p0, float *p1) const;
  void bindVelocity(const Velocity &_vel);
  void addSource(float *p, const float *source, const ShotPosition &pos) const;
  void addSource(float *p, const float *source, int is) const;
  void subSource(float *p, const float *source, const ShotPosition &pos) const;
  void addEncodedSource(float *p, const float *encsrc) const;
  void recordSeis(float *seis_it, const float *p) const;
  void maskGradient(float *grad) const;
  void scaleGradient(float *grad) const;
  void refillBoundary(float *vel) const;
  /* writes the physical part of exvel, recovered to m/s, as nz-float columns */
  void sfWriteVel(const std::vector<float> &exvel, FILE *file) const;

  void fwiRemoveDirectArrival(float *data, int shot_id) const;
  void removeDirectArrival(float *data) const;
  void subEncodedSource(float *p, const float *source) const;
  void refillVelStencilBndry();

  std::vector<float> initBndryVector(int nt) const;
  void writeBndry(float *_bndr, const float *p, int it) const;
  void readBndry(const float *_bndr, float *p, int it) const;

  void FwiForwardModeling(const std::vector<float> &encsrc, std::vector<float> &dcal, int shot_id) const;
  void EssForwardModeling(const std::vector<float> &encsrc, std::vector<float> &dcal) const;

  /* not in the reference: calgradient (file-local there, fwiframework.cpp:114
   * and essfwiframework.cpp:114) as a public call; shot_id < 0 = encoded */
  void calGradient(const std::vector<float> &encsrc, const std::vector<float> &vsrc, std::vector<float> &g0,
                   int shot_id) const;

 public:
  const Velocity &getVelocity() const;
  Velocity &getVelocity();
  const ShotPosition &getAllSrcPos() const;
  const ShotPosition &getAllGeoPos() const;
  int getns() const;
  int getng() const;
  float getdt() const;
  float getdx() const;
  int getnt() const;
  int getnx() const;
  int getnz() const;
  float getfm() const;
  int getnb() const;

  /* the device solver bound to the current velocity (uploaded on each call) */
  swfwi::Solver &device() const;

 private:
  void manipSource(float *p, const float *source, const ShotPosition &pos, int sign) const;
  void recordSeis(float *seis_it, const float *p, const ShotPosition &geoPos) const;
  void removeDirectArrival(const ShotPosition &allSrcPos, const ShotPosition &allGeoPos, float *data, int nt,
                           float t_width) const;
  void sync_velocity() const;

 private:
  static const int EXFDBNDRYLEN = 6;
  const Velocity *vel;
  const ShotPosition *allSrcPos;
  const ShotPosition *allGeoPos;
  float dt;
  float dx;
  float fm;
  int bx0, bxn;
  int bz0, bzn;
  int nt;
  int nb;
  mutable int bndrSize;
  mutable int bndrWidth;
  mutable std::shared_ptr<swfwi::Solver> solver_;
};

#endif /* SWFWI_DAMP4T10D_H_ */
