// Host-side pieces of the Damp4t10d surface that run once per model or per
// run: domain expansion, velocity transforms, the Ricker wavelet.  They keep
// the reference's float arithmetic so the expanded model and the wavelet are
// bit-identical to what the reference feeds its kernels.
#pragma once
#include <cmath>
#include <vector>

namespace swfwi {

constexpr int kExFd = 6;  // EXFDBNDRYLEN, damp4t10d.h:69

// expandBndry (damp4t10d.cpp:96-124), transvel (:126-130), expandForStencil /
// fillForStencil (:47-94): v [nx][nz] m/s -> out [nx+2nb+12][nz+nb+12].
inline void expand_domain(const float* v, int nx, int nz, int nb, float dx, float dt, float* out) {
  const int h = kExFd;
  const int nxb = nx + 2 * nb, nzb = nz + nb;
  std::vector<float> b((size_t)nxb * nzb, 0.f);
  for (int ix = 0; ix < nx; ix++)
    for (int iz = 0; iz < nz; iz++) b[(size_t)(nb + ix) * nzb + iz] = v[(size_t)ix * nz + iz];
  for (int ix = 0; ix < nb; ix++)
    for (int iz = 0; iz < nz; iz++) {
      b[(size_t)ix * nzb + iz] = v[iz];
      b[(size_t)(nb + nx + ix) * nzb + iz] = v[(size_t)(nx - 1) * nz + iz];
    }
  for (int ix = 0; ix < nxb; ix++)
    for (int iz = 0; iz < nb; iz++) b[(size_t)ix * nzb + (nz + iz)] = b[(size_t)ix * nzb + (nz - 1)];
  for (float& x : b) x = (dx * dx) / (x * x * dt * dt);
  const int nxp = nxb + 2 * h, nzp = nzb + 2 * h;
  for (int ix = h; ix < nxb + h; ix++)
    for (int iz = h; iz < nzb + h; iz++) out[(size_t)ix * nzp + iz] = b[(size_t)(ix - h) * nzb + (iz - h)];
  for (int ix = h; ix < nxb + h; ix++) {
    for (int iz = 0; iz < h; iz++) out[(size_t)ix * nzp + iz] = out[(size_t)ix * nzp + h];
    for (int iz = nzb + h; iz < nzp; iz++) out[(size_t)ix * nzp + iz] = out[(size_t)ix * nzp + (nzb + h - 1)];
  }
  for (int iz = 0; iz < nzp; iz++) {
    for (int ix = 0; ix < h; ix++) out[(size_t)ix * nzp + iz] = out[(size_t)h * nzp + iz];
    for (int ix = nxb + h; ix < nxp; ix++) out[(size_t)ix * nzp + iz] = out[(size_t)(nxb + h - 1) * nzp + iz];
  }
}

// recoverVel (damp4t10d.cpp:132-136)
inline float recover_vel(float vt, float dx, float dt) { return std::sqrt(dx * dx / (dt * dt * vt)); }
// velTrans / velRecover (common.h:21-30)
inline float vel_trans(float v, float dx, float dt) { return dx * dx / (dt * dt * v * v); }
inline float vel_recover(float vt, float dx, float dt) {
  float t = dx * dx / (dt * dt * vt);
  return std::sqrt(t);
}

// rickerWavelet (ricker-wavelet.cpp:15-21)
inline void ricker_wavelet(float* wlt, int nt, float fm, float dt, float amp) {
  for (int it = 0; it < nt; it++) {
    float tmp = M_PI * fm * (it * dt - 1.0 / fm);
    tmp *= tmp;
    wlt[it] = amp * (1.0 - 2.0 * tmp) * std::exp(-tmp);
  }
}

}  // namespace swfwi
