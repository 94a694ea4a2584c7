/* shot-position.h -- ShotPosition (shot-position.cpp:11-41 of swfwi): ns
 * positions pos = sz + nz*sx on a physical grid of nz rows. */
#ifndef SWFWI_SHOT_POSITION_H_
#define SWFWI_SHOT_POSITION_H_

#include <algorithm>
#include <vector>

class ShotPosition {
 public:
  ShotPosition(int szbeg, int sxbeg, int jsz, int jsx, int _ns, int _nz) : ns(_ns), pos(_ns), nz(_nz) {
    for (int is = 0; is < ns; is++) pos[is] = (szbeg + is * jsz) + nz * (sxbeg + is * jsx);
  }
  ShotPosition clipRange(int begin, int end) const {
    ShotPosition ret = *this;
    ret.ns = end - begin + 1;
    ret.pos.assign(pos.begin() + begin, pos.begin() + end + 1);
    return ret;
  }
  ShotPosition clip(int idx) const { return clipRange(idx, idx); }
  int getx(int idx) const { return pos[idx] / nz; }
  int getz(int idx) const { return pos[idx] % nz; }

 public:
  int ns;

 private:
  std::vector<int> pos;
  int nz;
};

#endif
