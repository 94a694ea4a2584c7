// The drop-in C++ classes (include/swfwi/*.h) over the device solver.
// Host-side loops keep the reference's arithmetic; time loops, gradients,
// epochs and analyses run in the device modules.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "../../include/swfwi/common.h"
#include "../../include/swfwi/damp4t10d.h"
#include "../../include/swfwi/enkfanalyze.h"
#include "../../include/swfwi/essfwiframework.h"
#include "../../include/swfwi/updatevelop.h"
#include "../../include/swfwi_hip.h"
#include "enkf.h"
#include "hostmath.h"
#include "solver.h"

// ---------------------------------------------------------------------------
// Damp4t10d

Damp4t10d::Damp4t10d(const ShotPosition& _allSrcPos, const ShotPosition& _allGeoPos, float _dt, float _dx, float _fm,
                     int _nb, int _nt)
    : vel(nullptr), allSrcPos(&_allSrcPos), allGeoPos(&_allGeoPos), dt(_dt), dx(_dx), fm(_fm), nt(_nt), nb(_nb),
      bndrSize(0), bndrWidth(6) {
  // damp4t10d.cpp:367-376
  bz0 = EXFDBNDRYLEN;
  bx0 = bxn = bzn = _nb + EXFDBNDRYLEN;
}

Damp4t10d::Damp4t10d(const Damp4t10d& o) = default;
Damp4t10d& Damp4t10d::operator=(const Damp4t10d& o) = default;
Damp4t10d::~Damp4t10d() = default;

Velocity Damp4t10d::expandDomain(const Velocity& _vel) {
  Velocity ret(_vel.nx + 2 * nb + 2 * EXFDBNDRYLEN, _vel.nz + nb + 2 * EXFDBNDRYLEN);
  swfwi::expand_domain(_vel.dat.data(), _vel.nx, _vel.nz, nb, dx, dt, ret.dat.data());
  return ret;
}

void Damp4t10d::bindVelocity(const Velocity& _vel) { vel = &_vel; }

void Damp4t10d::stepForward(float* p0, float* p1) const {
  std::vector<float> u2((size_t)vel->nx * vel->nz, 0.f);
  fd4t10s_damp_zjh_2d_vtrans(p0, p1, vel->dat.data(), u2.data(), vel->nx, vel->nz, bx0);
}

void Damp4t10d::stepBackward(float* p0, float* p1) const {
  std::vector<float> u2((size_t)vel->nx * vel->nz, 0.f);
  fd4t10s_zjh_2d_vtrans(p0, p1, vel->dat.data(), u2.data(), vel->nx, vel->nz);
}

void Damp4t10d::manipSource(float* p, const float* source, const ShotPosition& pos, int sign) const {
  // damp4t10d.cpp:204-214
  const int nzpad = vel->nz;
  for (int is = 0; is < pos.ns; is++) {
    const int sx = pos.getx(is) + bx0;
    const int sz = pos.getz(is) + bz0;
    float& v = p[(size_t)sx * nzpad + sz];
    v = sign > 0 ? v + source[is] : v - source[is];
  }
}

void Damp4t10d::addSource(float* p, const float* source, const ShotPosition& pos) const {
  manipSource(p, source, pos, +1);
}
void Damp4t10d::addSource(float*, const float*, int) const {}  // empty in the reference (damp4t10d.cpp:401-403)
void Damp4t10d::subSource(float* p, const float* source, const ShotPosition& pos) const {
  manipSource(p, source, pos, -1);
}
void Damp4t10d::addEncodedSource(float* p, const float* encsrc) const { addSource(p, encsrc, *allSrcPos); }
void Damp4t10d::subEncodedSource(float* p, const float* source) const { subSource(p, source, *allSrcPos); }

void Damp4t10d::recordSeis(float* seis_it, const float* p, const ShotPosition& geoPos) const {
  const int nzpad = vel->nz;
  for (int ig = 0; ig < geoPos.ns; ig++)
    seis_it[ig] = p[(size_t)(geoPos.getx(ig) + bx0) * nzpad + geoPos.getz(ig) + bz0];
}
void Damp4t10d::recordSeis(float* seis_it, const float* p) const { recordSeis(seis_it, p, *allGeoPos); }

void Damp4t10d::maskGradient(float* grad) const {
  const int nxpad = vel->nx, nzpad = vel->nz;
  for (int ix = 0; ix < nxpad; ix++)
    for (int iz = 0; iz < nzpad; iz++)
      if (ix < bx0 || iz < bz0 || ix >= nxpad - bxn || iz >= nzpad - bzn) grad[(size_t)ix * nzpad + iz] = 0.f;
}

void Damp4t10d::scaleGradient(float* grad) const {
  const int nxpad = vel->nx, nzpad = vel->nz;
  const int nz = nzpad - bz0 - bzn;
  for (int ix = bx0; ix < nxpad - bxn; ix++)
    for (int iz = 1; iz < nz; iz++) grad[(size_t)ix * nzpad + iz + bz0] *= std::sqrt(static_cast<float>(iz));
}

void Damp4t10d::refillBoundary(float* g) const {
  const int nzpad = vel->nz, nxpad = vel->nx;
  for (int ix = 0; ix < nxpad; ix++) {
    for (int iz = 0; iz < bz0; iz++) g[(size_t)ix * nzpad + iz] = g[(size_t)ix * nzpad + bz0];
    for (int iz = nzpad - bzn; iz < nzpad; iz++) g[(size_t)ix * nzpad + iz] = g[(size_t)ix * nzpad + nzpad - bzn - 1];
  }
  for (int ix = 0; ix < bx0; ix++)
    for (int iz = 0; iz < nzpad; iz++) g[(size_t)ix * nzpad + iz] = g[(size_t)bx0 * nzpad + iz];
  for (int ix = nxpad - bxn; ix < nxpad; ix++)
    for (int iz = 0; iz < nzpad; iz++) g[(size_t)ix * nzpad + iz] = g[(size_t)(nxpad - bxn - 1) * nzpad + iz];
}

void Damp4t10d::refillVelStencilBndry() {
  // fillForStencil on the bound model (damp4t10d.cpp:47-75): 6-cell halo
  Velocity& ev = getVelocity();
  const int h = EXFDBNDRYLEN, nxpad = ev.nx, nzpad = ev.nz;
  const int nx = nxpad - 2 * h, nz = nzpad - 2 * h;
  float* v = ev.dat.data();
  for (int ix = h; ix < nx + h; ix++) {
    for (int iz = 0; iz < h; iz++) v[(size_t)ix * nzpad + iz] = v[(size_t)ix * nzpad + h];
    for (int iz = nz + h; iz < nzpad; iz++) v[(size_t)ix * nzpad + iz] = v[(size_t)ix * nzpad + (nz + h - 1)];
  }
  for (int iz = 0; iz < nzpad; iz++) {
    for (int ix = 0; ix < h; ix++) v[(size_t)ix * nzpad + iz] = v[(size_t)h * nzpad + iz];
    for (int ix = nx + h; ix < nxpad; ix++) v[(size_t)ix * nzpad + iz] = v[(size_t)(nx + h - 1) * nzpad + iz];
  }
}

void Damp4t10d::sfWriteVel(const std::vector<float>& exvel, FILE* file) const {
  // damp4t10d.cpp:254-265: recoverVel, then the physical columns
  const int nzpad = vel->nz, nxpad = vel->nx;
  const int nz = nzpad - bz0 - bzn;
  std::vector<float> vv = exvel;
  for (float& x : vv) x = std::sqrt(dx * dx / (dt * dt * x));
  for (int ix = bx0; ix < nxpad - bxn; ix++)
    if (std::fwrite(&vv[(size_t)ix * nzpad + EXFDBNDRYLEN], sizeof(float), nz, file) != (size_t)nz)
      throw std::runtime_error("sfWriteVel: short write");
}

void Damp4t10d::removeDirectArrival(const ShotPosition& srcPos, const ShotPosition& geoPos, float* data, int ntt,
                                    float t_width) const {
  // damp4t10d.cpp:321-365 (the transposes only reorder; the mute is per trace)
  const int half_len = t_width / dt;
  const int sx = srcPos.getx(0) + bx0;
  const int sz = srcPos.getz(0) + bz0;
  const int gz0 = geoPos.getz(0) + bz0;
  float vel_average = 0.0;
  const int gmin = (sz < gz0) ? sz : gz0, gmax = (sz > gz0) ? sz : gz0;
  const std::vector<float>& vv = vel->dat;
  const int nx = vel->nx, nz = vel->nz;
  for (int i = 0; i < nx; i++)
    for (int k = gmin; k <= gmax; k++) vel_average += vv[(size_t)i * nz + k];
  vel_average /= nx * (gmax - gmin + 1);
  const int ng = geoPos.ns;
  for (int itr = 0; itr < ng; itr++) {
    const int gx = geoPos.getx(itr) + bx0, gz = geoPos.getz(itr) + bz0;
    const float dist = (gx - sx) * (gx - sx) + (gz - sz) * (gz - sz);
    const int t = (int)std::sqrt(dist * vel_average);
    const int end = ((t + 2 * half_len) > ntt) ? ntt : (t + 2 * half_len);
    for (int j = t; j < end; j++) data[(size_t)j * ng + itr] = 0.f;
  }
}

void Damp4t10d::fwiRemoveDirectArrival(float* data, int shot_id) const {
  const float t_width = 1.5 / fm;
  const ShotPosition cur = allSrcPos->clipRange(shot_id, shot_id);
  removeDirectArrival(cur, *allGeoPos, data, nt, t_width);
}

void Damp4t10d::removeDirectArrival(float* data) const {
  const float t_width = 1.5 / fm;
  removeDirectArrival(*allSrcPos, *allGeoPos, data, nt, t_width);
}

std::vector<float> Damp4t10d::initBndryVector(int ntt) const {
  if (vel == nullptr) {
    std::fprintf(stderr, "initBndryVector: you should bind velocity first\n");
    std::exit(1);
  }
  const int nxpad = vel->nx, nzpad = vel->nz;
  bndrWidth = 6;
  const int nx = nxpad - (bx0 - bndrWidth + bxn - bndrWidth);
  const int nz = nzpad - bz0 - bzn;
  bndrSize = bndrWidth * (nx + 2 * nz);
  return std::vector<float>((size_t)ntt * bndrSize, 0);
}

void Damp4t10d::writeBndry(float* _bndr, const float* p, int it) const {
  const int nxpad = vel->nx, nzpad = vel->nz;
  const int nx = nxpad - (bx0 - bndrWidth + bxn - bndrWidth);
  const int nz = nzpad - bz0 - bzn;
  float* b = &_bndr[(size_t)it * bndrSize];
  for (int ix = 0; ix < nx; ix++)
    for (int iz = 0; iz < bndrWidth; iz++)
      b[iz + bndrWidth * ix] = p[(size_t)(ix + bx0 - bndrWidth) * nzpad + (nzpad - bzn + iz)];
  for (int iz = 0; iz < nz; iz++)
    for (int ix = 0; ix < bndrWidth; ix++) {
      b[bndrWidth * nx + iz + nz * ix] = p[(size_t)(bx0 - bndrWidth + ix) * nzpad + (bz0 + iz)];
      b[bndrWidth * nx + iz + nz * (ix + bndrWidth)] = p[(size_t)(nxpad - bxn + ix) * nzpad + (bz0 + iz)];
    }
}

void Damp4t10d::readBndry(const float* _bndr, float* p, int it) const {
  const int nxpad = vel->nx, nzpad = vel->nz;
  const int nx = nxpad - (bx0 - bndrWidth + bxn - bndrWidth);
  const int nz = nzpad - bz0 - bzn;
  const float* b = &_bndr[(size_t)it * bndrSize];
  for (int ix = 0; ix < nx; ix++)
    for (int iz = 0; iz < bndrWidth; iz++)
      p[(size_t)(ix + bx0 - bndrWidth) * nzpad + (nzpad - bzn + iz)] = b[iz + bndrWidth * ix];
  for (int iz = 0; iz < nz; iz++)
    for (int ix = 0; ix < bndrWidth; ix++) {
      p[(size_t)(bx0 - bndrWidth + ix) * nzpad + (bz0 + iz)] = b[bndrWidth * nx + iz + nz * ix];
      p[(size_t)(nxpad - bxn + ix) * nzpad + (bz0 + iz)] = b[bndrWidth * nx + iz + nz * (ix + bndrWidth)];
    }
}

swfwi::Solver& Damp4t10d::device() const {
  if (vel == nullptr) throw std::runtime_error("Damp4t10d: bind a velocity before running on the device");
  if (!solver_) {
    swfwi::Geometry g;
    g.nx = vel->nx - 2 * nb - 2 * EXFDBNDRYLEN;
    g.nz = vel->nz - nb - 2 * EXFDBNDRYLEN;
    g.nb = nb;
    g.dt = dt;
    g.dx = dx;
    g.fm = fm;
    g.nt = nt;
    for (int i = 0; i < allSrcPos->ns; i++) {
      g.sx.push_back(allSrcPos->getx(i));
      g.sz.push_back(allSrcPos->getz(i));
    }
    for (int i = 0; i < allGeoPos->ns; i++) {
      g.gx.push_back(allGeoPos->getx(i));
      g.gz.push_back(allGeoPos->getz(i));
    }
    solver_ = std::make_shared<swfwi::Solver>(g, swfwi::kExact);
  }
  return *solver_;
}

void Damp4t10d::sync_velocity() const { device().bind_velocity(vel->dat.data()); }

void Damp4t10d::FwiForwardModeling(const std::vector<float>& encsrc, std::vector<float>& dcal, int shot_id) const {
  sync_velocity();
  dcal.resize((size_t)nt * getng());
  device().forward_modeling(encsrc.data(), shot_id, dcal.data());
}

void Damp4t10d::EssForwardModeling(const std::vector<float>& encsrc, std::vector<float>& dcal) const {
  sync_velocity();
  dcal.resize((size_t)nt * getng());
  device().forward_modeling(encsrc.data(), -1, dcal.data());
}

void Damp4t10d::calGradient(const std::vector<float>& encsrc, const std::vector<float>& vsrc, std::vector<float>& g0,
                            int shot_id) const {
  sync_velocity();
  std::vector<float> g((size_t)vel->nx * vel->nz);
  device().calgradient(encsrc.data(), shot_id, vsrc.data(), g.data());
  g0.resize(g.size(), 0.f);
  for (size_t i = 0; i < g.size(); i++) g0[i] = g0[i] + g[i];
}

const Velocity& Damp4t10d::getVelocity() const { return *vel; }
Velocity& Damp4t10d::getVelocity() { return *const_cast<Velocity*>(vel); }
const ShotPosition& Damp4t10d::getAllSrcPos() const { return *allSrcPos; }
const ShotPosition& Damp4t10d::getAllGeoPos() const { return *allGeoPos; }
int Damp4t10d::getns() const { return allSrcPos->ns; }
int Damp4t10d::getng() const { return allGeoPos->ns; }
float Damp4t10d::getdt() const { return dt; }
float Damp4t10d::getdx() const { return dx; }
int Damp4t10d::getnt() const { return nt; }
int Damp4t10d::getnx() const { return vel->nx; }
int Damp4t10d::getnz() const { return vel->nz; }
float Damp4t10d::getfm() const { return fm; }
int Damp4t10d::getnb() const { return nb; }

// ---------------------------------------------------------------------------
// UpdateVelOp

UpdateVelOp::UpdateVelOp(float _vmin, float _vmax, float dx, float dt) : vmin_ms(_vmin), vmax_ms(_vmax) {
  // updatevelop.cpp:11-21 of the listing
  vmin = (dx / dt / _vmax) * (dx / dt / _vmax);
  vmax = (dx / dt / _vmin) * (dx / dt / _vmin);
  if (vmax <= vmin) {
    std::fprintf(stderr, "vmax(%f) < _vmin(%f)\n", vmax, vmin);
    std::exit(0);
  }
}

void UpdateVelOp::update(Velocity& newVel, const Velocity& v, const std::vector<float>& grad, float steplen) const {
  const size_t n = newVel.dat.size();
  for (size_t i = 0; i < n; i++) {
    float x = v.dat[i] + steplen * grad[i];
    if (x > vmax) x = vmax;
    if (x < vmin) x = vmin;
    newVel.dat[i] = x;
  }
}

// ---------------------------------------------------------------------------
// EssFwiFramework / FwiFramework

EssFwiFramework::EssFwiFramework(Damp4t10d& method, const UpdateSteplenOp& steplen, const UpdateVelOp& velop,
                                 const std::vector<float>& wlt, const std::vector<float>& dobs)
    : fmMethod(method), ess(true) {
  swfwi::UpdateParams p;
  p.vmin = velop.vminPhysical();
  p.vmax = velop.vmaxPhysical();
  p.nita = steplen.maxIterSelectAlpha3();
  p.maxdv = steplen.maxDv();
  inv.reset(new swfwi::Inversion(method.device(), wlt.data(), dobs.data(), p));
}

EssFwiFramework::~EssFwiFramework() = default;

void EssFwiFramework::epoch(int iter) {
  Velocity& v = fmMethod.getVelocity();
  fmMethod.device().bind_velocity(v.dat.data());
  if (ess)
    inv->ess_epoch(iter);
  else
    inv->fwi_epoch(iter);
  fmMethod.device().download_velocity(v.dat.data());
}

void EssFwiFramework::writeVel(FILE* file) const { fmMethod.sfWriteVel(fmMethod.getVelocity().dat, file); }
float EssFwiFramework::getUpdateObj() const { return inv->update_obj(); }
float EssFwiFramework::getInitObj() const { return inv->init_obj(); }

FwiFramework::FwiFramework(Damp4t10d& method, const UpdateSteplenOp& steplen, const UpdateVelOp& velop,
                           const std::vector<float>& wlt, const std::vector<float>& dobs)
    : EssFwiFramework(method, steplen, velop, wlt, dobs) {
  ess = false;
}

// ---------------------------------------------------------------------------
// EnkfAnalyze

EnkfAnalyze::EnkfAnalyze(const Damp4t10d& _fm, const std::vector<float>& _wlt, const std::vector<float>& _dobs,
                         float sigmafactor)
    : fm(_fm), wlt(_wlt), dobs(_dobs), sigmaFactor(sigmafactor) {}

EnkfAnalyze::~EnkfAnalyze() = default;

void EnkfAnalyze::pAnalyze(std::vector<float*>& velSet) const {
  const int N = (int)velSet.size();
  swfwi::Solver& s = fm.device();
  const size_t n = (size_t)s.npts();
  if (!impl) impl.reset(new swfwi::Enkf(s, wlt.data(), dobs.data(), sigmaFactor, N));
  if (impl->members() != N) throw std::runtime_error("EnkfAnalyze: ensemble size changed between analyses");
  std::vector<float> all(n * N);
  for (int i = 0; i < N; i++) std::memcpy(&all[n * i], velSet[i], n * sizeof(float));
  impl->set_members(all.data());
  impl->analyze(nullptr, seed_ + calls_++);
  impl->get_members(all.data());
  for (int i = 0; i < N; i++) std::memcpy(velSet[i], &all[n * i], n * sizeof(float));
}

std::vector<float> EnkfAnalyze::pCreateAMean(const std::vector<float*>& velSet, const int nSamples) const {
  // enkfanalyze.cpp:578-597 on one process
  const size_t n = (size_t)fm.getnx() * fm.getnz();
  std::vector<float> ret(n, 0.f);
  for (size_t i = 0; i < n; i++)
    for (size_t j = 0; j < velSet.size(); j++) ret[i] += velSet[j][i];
  for (size_t i = 0; i < n; i++) ret[i] /= nSamples;
  return ret;
}

// ---------------------------------------------------------------------------
// common helpers

void matrix_transpose(float* matrix, float* trans, int n1, int n2) {
  for (int i2 = 0; i2 < n2; i2++)
    for (int i1 = 0; i1 < n1; i1++) trans[i2 + n2 * i1] = matrix[i1 + n1 * i2];
}

float cal_objective(float* dres, int ng) {
  float obj = 0;
  for (int i = 0; i < ng; i++) {
    float a = dres[i];
    obj += a * a;
  }
  return obj;
}

void rickerWavelet(float* wlt, int nt, float fm, float dt, float amp) { swfwi::ricker_wavelet(wlt, nt, fm, dt, amp); }

std::vector<float> rickerWavelet(int nt, float fm, float dt, float amp) {
  std::vector<float> w(nt);
  swfwi::ricker_wavelet(w.data(), nt, fm, dt, amp);
  return w;
}

void parabolaVertex(float x1, float y1, float x2, float y2, float x3, float y3, float, float& xv, float& yv) {
  double k2 = (y3 - y2) / (x3 - x2);
  double k1 = (y2 - y1) / (x2 - x1);
  double denom = (x1 - x2) * (x1 - x3) * (x2 - x3);
  double A = (x3 * (y2 - y1) + x2 * (y1 - y3) + x1 * (y3 - y2)) / denom;
  double B = (x3 * x3 * (y1 - y2) + x2 * x2 * (y3 - y1) + x1 * x1 * (y2 - y3)) / denom;
  double C = (x2 * x3 * (x2 - x3) * y1 + x3 * x1 * (x3 - x1) * y2 + x1 * x2 * (x1 - x2) * y3) / denom;
  xv = -B / (2 * A);
  yv = C - B * B / (4 * A);
  if (std::abs(k2 - k1) < 0.001 * (std::max(std::abs(k2), std::abs(k1))) || std::isnan(-xv)) {
    xv = x3;
    yv = y3;
  }
}

RandomCodes::RandomCodes(int seed) {
  uint64_t s = (uint64_t)(uint32_t)seed % 2147483647u;
  x_ = s == 0 ? 1 : s;
}

int RandomCodes::nextRand() {
  x_ = (x_ * 48271u) % 2147483647u;
  return (int)((x_ - 1u) / 1073741823u) * 2 - 1;
}

std::vector<int> RandomCodes::genPlus1Minus1(int nshots) {
  std::vector<int> c(nshots);
  for (int i = 0; i < nshots; i++) c[i] = nextRand();
  return c;
}

std::vector<float> Encoder::encodeSource(const std::vector<float>& wlt) {
  const int ns = (int)mCode.size(), nt = (int)wlt.size();
  std::vector<float> enc((size_t)ns * nt);
  for (int it = 0; it < nt; it++)
    for (int is = 0; is < ns; is++) enc[(size_t)it * ns + is] = wlt[it] * mCode[is];
  return enc;
}

std::vector<float> Encoder::encodeObsData(const std::vector<float>& dobs, int nt, int ng) {
  std::vector<float> enc((size_t)nt * ng, 0);
  const int ns = (int)mCode.size();
  const size_t m = (size_t)nt * ng;
  for (int is = 0; is < ns; is++)
    for (size_t j = 0; j < m; j++) enc[j] += dobs[is * m + j] * mCode[is];
  return enc;
}
