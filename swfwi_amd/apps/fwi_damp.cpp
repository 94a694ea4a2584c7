// essfwi-damp / fwi-damp: encoded-source or per-shot FWI iterations
// (main-essfwi-damp.cpp:145-211, main-fwi-damp.cpp:145-212 of swfwi) with the
// same parameters, inputs and outputs (vout, absobjs, norobjs).  Built twice:
// FWI_PER_SHOT selects FwiFramework.
#include <cstdio>
#include <vector>

#include "../../include/swfwi/common.h"
#include "../../include/swfwi/damp4t10d.h"
#include "../../include/swfwi/essfwiframework.h"
#include "../../include/swfwi/updatesteplenop.h"
#include "../../include/swfwi/updatevelop.h"
#include "rsfio.h"

int main(int argc, char** argv) {
  rsf::Params par(argc, argv);
  rsf::Input vinit(par.need("vin"));
  rsf::Input shots(par.need("shots"));
  int niter, nita, seed, nz, nx, nt, ng, ns, sxbeg, szbeg, gxbeg, gzbeg, jsx, jsz, jgx, jgz, nb;
  float maxdv, dz, dx, dt, amp, fm, vmin, vmax;
  if (!par.getint("niter", niter)) rsf::die("no niter");
  if (!par.getfloat("maxdv", maxdv)) rsf::die("no maxdv");
  if (!par.getint("nita", nita)) rsf::die("no nita");
  if (!par.getint("seed", seed)) seed = 10;
  if (!vinit.histint("n1", nz)) rsf::die("no n1");
  if (!vinit.histint("n2", nx)) rsf::die("no n2");
  if (!vinit.histfloat("d1", dz)) rsf::die("no d1");
  if (!vinit.histfloat("d2", dx)) rsf::die("no d2");
  if (!shots.histint("n1", nt)) rsf::die("no nt");
  if (!shots.histint("n2", ng)) rsf::die("no ng");
  if (!shots.histint("n3", ns)) rsf::die("no ns");
  if (!shots.histfloat("d1", dt)) rsf::die("no dt");
  if (!shots.histfloat("amp", amp)) rsf::die("no amp");
  if (!shots.histfloat("fm", fm)) rsf::die("no fm");
  if (!shots.histint("sxbeg", sxbeg)) rsf::die("no sxbeg");
  if (!shots.histint("szbeg", szbeg)) rsf::die("no szbeg");
  if (!shots.histint("gxbeg", gxbeg)) rsf::die("no gxbeg");
  if (!shots.histint("gzbeg", gzbeg)) rsf::die("no gzbeg");
  if (!shots.histint("jsx", jsx)) rsf::die("no jsx");
  if (!shots.histint("jsz", jsz)) rsf::die("no jsz");
  if (!shots.histint("jgx", jgx)) rsf::die("no jgx");
  if (!shots.histint("jgz", jgz)) rsf::die("no jgz");
  if (!shots.histint("nb", nb)) rsf::die("no nb");
  if (!shots.histfloat("vmin", vmin)) rsf::die("no vmin");
  if (!shots.histfloat("vmax", vmax)) rsf::die("no vmax");
  (void)seed;

  rsf::Output vout(par.need("vout"));
  rsf::Output absobjs(par.need("absobjs"));
  rsf::Output norobjs(par.need("norobjs"));
  vout.putint("n1", nz);
  vout.putint("n2", nx);
  vout.putfloat("d1", dz);
  vout.putfloat("d2", dx);
  vout.putstring("label1", "Depth");
  vout.putstring("label2", "Distance");
  vout.putstring("label3", "Iteration");
  vout.putint("n3", niter);
  vout.putint("d3", 1);
  vout.putint("o3", 1);
  for (rsf::Output* o : {&absobjs, &norobjs}) {
    o->putint("n1", niter + 1);
    o->putfloat("d1", 1);
    o->putfloat("o1", 1);
  }
  absobjs.putstring("label1", "Absolute");
  norobjs.putstring("label1", "Normalize");

  ShotPosition allSrcPos(szbeg, sxbeg, jsz, jsx, ns, nz);
  ShotPosition allGeoPos(gzbeg, gxbeg, jgz, jgx, ng, nz);
  Damp4t10d fmMethod(allSrcPos, allGeoPos, dt, dx, fm, nb, nt);
  Velocity v0(vinit.floats((size_t)nx * nz), nx, nz);
  Velocity exvel = fmMethod.expandDomain(v0);
  fmMethod.bindVelocity(exvel);
  std::vector<float> wlt(nt);
  rickerWavelet(&wlt[0], nt, fm, dt, amp);

  // ShotDataReader::serialRead (shotdata-reader.cpp:54-64): [ns][ng][nt] on
  // disk, [ns][nt][ng] in memory
  std::vector<float> disk = shots.floats((size_t)ns * ng * nt);
  std::vector<float> dobs((size_t)ns * nt * ng);
  for (int is = 0; is < ns; is++) matrix_transpose(&disk[(size_t)is * ng * nt], &dobs[(size_t)is * nt * ng], nt, ng);

  UpdateVelOp updatevelop(vmin, vmax, dx, dt);
  UpdateSteplenOp updateSteplenOp(fmMethod, updatevelop, nita, maxdv);
#ifdef FWI_PER_SHOT
  FwiFramework fwi(fmMethod, updateSteplenOp, updatevelop, wlt, dobs);
#else
  EssFwiFramework fwi(fmMethod, updateSteplenOp, updatevelop, wlt, dobs);
#endif
  std::vector<float> absobj, norobj;
  for (int iter = 0; iter < niter; iter++) {
    fwi.epoch(iter);
    fwi.writeVel(vout.data());
    const float obj = fwi.getUpdateObj();
    if (iter == 0) {
      absobj.push_back(fwi.getInitObj());
      norobj.push_back(1);
    }
    absobj.push_back(obj);
    norobj.push_back(obj / absobj[0]);
    std::printf("iter %d obj %e normalized %f\n", iter, obj, norobj.back());
  }
  absobjs.floatwrite(absobj.data(), absobj.size());
  norobjs.floatwrite(norobj.data(), norobj.size());
  return 0;
}
