/* common.h -- helpers of the swfwi hot path (common.h/common.cpp,
 * ricker-wavelet.cpp, random-code.cpp, encoder.cpp, parabola-vertex.cpp of
 * swfwi) with the reference's float/double arithmetic. */
#ifndef SWFWI_COMMON_H_
#define SWFWI_COMMON_H_

#include <cmath>
#include <cstdint>
#include <functional>
#include <vector>
#include <algorithm>

template <typename T>
void vectorMinus(const std::vector<T> &dobs, const std::vector<T> &dcal, std::vector<T> &vsrc) {
  std::transform(dobs.begin(), dobs.end(), dcal.begin(), vsrc.begin(), std::minus<T>());
}
template <typename E>
float velRecover(float vel, float dx, float dt) {
  float t = dx * dx / (dt * dt * vel);
  return std::sqrt(t);
}
template <typename E>
float velTrans(float vel, float dx, float dt) {
  return dx * dx / (dt * dt * vel * vel);
}

void matrix_transpose(float *matrix, float *trans, int n1, int n2);
float cal_objective(float *dres, int ng);

void rickerWavelet(float *wlt, int nt, float fm, float dt, float amp);
std::vector<float> rickerWavelet(int nt, float fm, float dt, float amp);

void parabolaVertex(float x1, float y1, float x2, float y2, float x3, float y3, float max_alpha3, float &xv, float &yv);

/* RandomCodes (random-code.h:18-20): minstd_rand + uniform_int<>(0,1) -> +-1 */
class RandomCodes {
 public:
  explicit RandomCodes(int seed);
  std::vector<int> genPlus1Minus1(int nshots);
  int nextRand();

 private:
  uint64_t x_;
};

/* Encoder (encoder.h:13-21) */
class Encoder {
 public:
  explicit Encoder(const std::vector<int> &code) : mCode(code) {}
  std::vector<float> encodeSource(const std::vector<float> &wlt);
  std::vector<float> encodeObsData(const std::vector<float> &dobs, int nt, int ng);

 private:
  const std::vector<int> &mCode;
};

#endif
