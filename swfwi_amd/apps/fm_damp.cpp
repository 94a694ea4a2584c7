// fm-damp: observed shots from a velocity model (main-fm-damp.cpp:202-282 of
// swfwi).  Same parameters and output header; every shot runs in one batched
// device launch per time step instead of one MPI rank per shot block.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <vector>

#include "../../include/swfwi/common.h"
#include "../../include/swfwi/damp4t10d.h"
#include "../csrc/solver.h"
#include "rsfio.h"

int main(int argc, char** argv) {
  rsf::Params par(argc, argv);
  rsf::Input vinit(par.need("vinit"));
  int nz, nx, nb, nt, ns, ng, jsx, jsz, jgx, jgz, sxbeg, szbeg, gxbeg, gzbeg;
  float dz, dx, dt, amp, fm;
  if (!vinit.histint("n1", nz)) rsf::die("no n1");
  if (!vinit.histint("n2", nx)) rsf::die("no n2");
  if (!vinit.histfloat("d1", dz)) rsf::die("no d1");
  if (!vinit.histfloat("d2", dx)) rsf::die("no d2");
  if (!par.getfloat("amp", amp)) amp = 1000;
  if (!par.getfloat("fm", fm)) fm = 10;
  if (!par.getint("nb", nb)) nb = 30;
  if (!par.getfloat("dt", dt)) rsf::die("no dt");
  if (!par.getint("nt", nt)) rsf::die("no nt");
  if (!par.getint("ns", ns)) rsf::die("no ns");
  if (!par.getint("ng", ng)) rsf::die("no ng");
  if (!par.getint("jsx", jsx)) rsf::die("no jsx");
  if (!par.getint("jsz", jsz)) jsz = 0;
  if (!par.getint("jgx", jgx)) jgx = 1;
  if (!par.getint("jgz", jgz)) jgz = 0;
  if (!par.getint("sxbeg", sxbeg)) rsf::die("no sxbeg");
  if (!par.getint("szbeg", szbeg)) rsf::die("no szbeg");
  if (!par.getint("gxbeg", gxbeg)) rsf::die("no gxbeg");
  if (!par.getint("gzbeg", gzbeg)) rsf::die("no gzbeg");
  // Params::check (main-fm-damp.cpp:183-194)
  if (!(sxbeg >= 0 && szbeg >= 0 && sxbeg + (ns - 1) * jsx < nx && szbeg + (ns - 1) * jsz < nz)) {
    std::fprintf(stderr, "sources exceeds the computing zone!\n");
    return 1;
  }
  if (!(gxbeg >= 0 && gzbeg >= 0 && gxbeg + (ng - 1) * jgx < nx && gzbeg + (ng - 1) * jgz < nz)) {
    std::fprintf(stderr, "geophones exceeds the computing zone!\n");
    return 1;
  }
  Velocity v(vinit.floats((size_t)nx * nz), nx, nz);
  const float vmin = *std::min_element(v.dat.begin(), v.dat.end());
  const float vmax = *std::max_element(v.dat.begin(), v.dat.end());

  rsf::Output shots(par.need("shots"));
  shots.putint("n1", nt);
  shots.putint("n2", ng);
  shots.putint("n3", ns);
  shots.putfloat("d1", dt);
  shots.putfloat("d2", jgx * dx);
  shots.putfloat("o1", 0);
  shots.putfloat("o2", 0);
  shots.putstring("label1", "Time");
  shots.putstring("label2", "Lateral");
  shots.putstring("label3", "Shot");
  shots.putstring("unit1", "sec");
  shots.putstring("unit2", "m");
  shots.putfloat("amp", amp);
  shots.putfloat("fm", fm);
  shots.putint("ng", ng);
  shots.putint("szbeg", szbeg);
  shots.putint("sxbeg", sxbeg);
  shots.putint("gzbeg", gzbeg);
  shots.putint("gxbeg", gxbeg);
  shots.putint("jsx", jsx);
  shots.putint("jsz", jsz);
  shots.putint("jgx", jgx);
  shots.putint("jgz", jgz);
  shots.putint("nb", nb);
  shots.putfloat("vmin", vmin);
  shots.putfloat("vmax", vmax);

  const auto t0 = std::chrono::steady_clock::now();
  ShotPosition allSrcPos(szbeg, sxbeg, jsz, jsx, ns, nz);
  ShotPosition allGeoPos(gzbeg, gxbeg, jgz, jgx, ng, nz);
  Damp4t10d fmMethod(allSrcPos, allGeoPos, dt, dx, fm, nb, nt);
  Velocity exvel = fmMethod.expandDomain(v);
  fmMethod.bindVelocity(exvel);
  std::vector<float> wlt(nt);
  rickerWavelet(&wlt[0], nt, fm, dt, amp);

  std::vector<float> dobs((size_t)ns * nt * ng);
  swfwi::Solver& dev = fmMethod.device();
  dev.bind_velocity(exvel.dat.data());
  dev.forward_shots(wlt.data(), dobs.data());
  std::vector<float> trans((size_t)nt * ng);
  for (int is = 0; is < ns; is++) {
    matrix_transpose(&dobs[(size_t)is * nt * ng], &trans[0], ng, nt);
    shots.floatwrite(trans.data(), trans.size());
  }
  shots.close();
  const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("fm-damp: %d shots, %d steps, total elapsed time %fs\n", ns, nt, el);
  return 0;
}
