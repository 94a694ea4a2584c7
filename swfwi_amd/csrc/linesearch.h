// Step-length search of UpdateSteplenOp (updatesteplenop.cpp:88-285) and the
// FP64 parabola vertex (parabola-vertex.cpp:12-33), as host-side state that
// any number of models can run side by side: each model's search asks for
// the objectives of a few trial step lengths at a time, the caller evaluates
// the requests of all models in one batched forward run, and feeds them back.
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

namespace swfwi {

inline float minf_std(float a, float b) { return (b < a) ? b : a; }  // std::min

// begin() of the parabolicLessComp-ordered std::set of (alpha, obj)
// (updatesteplenop.cpp:19-21,107): an insert becomes the first element when
// its objective is below the current first one's + 1e-10.
struct BestPoint {
  float a = 0, o = 0;
  bool any = false;
  void insert(float na, float no) {
    if (!any || (no - o) < 1e-10) {
      a = na;
      o = no;
      any = true;
    }
  }
};

// preservedAlpha + initAlpha23 (updatesteplenop.cpp:273-285).
struct PreservedAlpha {
  float alpha = 0;
  bool init = false;
  void init_alpha23(float maxAlpha3, float& a2, float& a3) {
    const float minAlpha = 1.0E-7, resetAlpha = 1.0E-4;
    if (!init) {
      init = true;
      alpha = maxAlpha3;
    }
    a3 = alpha;
    a3 = a3 < minAlpha ? resetAlpha : a3;
    a2 = a3 * 0.5;
  }
};

// parabolaVertex in FP64 with the degeneracy fallback to (x3, y3).
inline void parabola_vertex(float x1, float y1, float x2, float y2, float x3, float y3, float& xv, float& yv) {
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

// refineAlpha as written (updatesteplenop.cpp:88-214), including the two
// fallbacks that evaluate the loop's local alpha3 / alpha2 rather than the
// returned ones.  Driving loop:
//     RefineAlpha r(obj1, max3, a2, a3, nita);
//     while (!r.done()) r.feed(objectives_of(r.pending()));
class RefineAlpha {
 public:
  RefineAlpha() = default;
  RefineAlpha(float obj1, float maxAlpha3, float alpha2, float alpha3, int nita)
      : obj1_(obj1), max3_(maxAlpha3), nita_(nita), a2_(alpha2), a3_(alpha3) {
    pend_ = {a2_, a3_};
    st_ = kFirst;
  }
  bool done() const { return st_ == kDone; }
  const std::vector<float>& pending() const { return pend_; }
  // objectives of pending(), in order
  void feed(const float* obj) {
    switch (st_) {
      case kFirst:
        o2_ = obj[0];
        o3_ = obj[1];
        tuned_.insert(a2_, o2_);
        iter_ = 0;
        halve_or_next();
        break;
      case kHalve:
        o2_ = obj[0];
        tuned_.insert(a2_, o2_);
        iter_++;
        halve_or_next();
        break;
      case kHalveFallback:
        ro3 = obj[0];
        finish(0);
        break;
      case kDouble:
        o3_ = obj[0];
        tuned_.insert(a3_, o3_);
        double_or_next();
        break;
      case kDoubleFallback:
        ro2 = obj[0];
        finish(0);
        break;
      case kDone:
        break;
    }
  }
  // results (valid once done()): the bracket and whether to fit a parabola
  float ra2 = 0, ro2 = 0, ra3 = 0, ro3 = 0;
  int parabolic = 0;

 private:
  enum State { kFirst, kHalve, kHalveFallback, kDouble, kDoubleFallback, kDone };
  void halve_or_next() {
    if (iter_ < nita_ && o2_ > obj1_) {  // shrink alpha2 until it beats alpha1
      a3_ = a2_;
      o3_ = o2_;
      a2_ /= 2;
      pend_ = {a2_};
      st_ = kHalve;
      return;
    }
    if (o2_ > obj1_) {
      ra2 = tuned_.a;
      ro2 = tuned_.o;
      ra3 = minf_std(ra2 * 2, max3_);
      pend_ = {a3_};
      st_ = kHalveFallback;
      return;
    }
    const float alpha1 = 0;
    lin3_ = (o2_ - obj1_) / (a2_ - alpha1) * (a3_ - alpha1) + obj1_;
    tuned_ = BestPoint();
    tuned_.insert(a3_, o3_);
    double_or_next();
  }
  void double_or_next() {
    if (o3_ < lin3_ && o3_ < obj1_ && a3_ < max3_) {  // grow alpha3 while it keeps improving
      a2_ = a3_;
      o2_ = o3_;
      a3_ = minf_std(a3_ * 2, max3_);
      pend_ = {a3_};
      st_ = kDouble;
      return;
    }
    if (a3_ > max3_ + 0.1) {
      ra3 = tuned_.a;
      ro3 = tuned_.o;
      ra2 = ra3 / 2;
      pend_ = {a2_};
      st_ = kDoubleFallback;
      return;
    }
    ra2 = a2_;
    ra3 = a3_;
    ro2 = o2_;
    ro3 = o3_;
    finish(1);
  }
  void finish(int parab) {
    parabolic = parab;
    pend_.clear();
    st_ = kDone;
  }

  float obj1_ = 0, max3_ = 0;
  int nita_ = 0, iter_ = 0;
  float a2_ = 0, a3_ = 0, o2_ = 0, o3_ = 0, lin3_ = 0;
  BestPoint tuned_;
  std::vector<float> pend_;
  State st_ = kDone;
};

// calsteplen's ending (updatesteplenop.cpp:239-264): the step and its
// objective from a finished search.
inline void final_step(const RefineAlpha& r, float obj1, float maxAlpha3, float& alpha4, float& obj4) {
  if (r.parabolic) {
    parabola_vertex(0, obj1, r.ra2, r.ro2, r.ra3, r.ro3, alpha4, obj4);
    if (alpha4 > maxAlpha3) alpha4 = maxAlpha3;
  } else {
    alpha4 = r.ra3;
    obj4 = r.ro3;
  }
}

}  // namespace swfwi
