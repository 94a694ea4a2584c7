// enfwi-damp: ensemble-Kalman FWI (main-enfwi-damp.cpp:406-631 of swfwi) with
// the same parameters, inputs and outputs, on one GPU through the C ABI:
// every member's ESS-FWI epoch runs as one batched epoch, the EnKF analysis
// every `niterenkf` iterations, the mean model's objective after each.
//
//   vin= vreal= shots= perin= vout= absobjs= norobjs= niter= maxdv= nita=
//   nsample= niterenkf= sigfac= [seed=10]
//
// perin holds nsample raw float32 perturbations (m/s) of the padded model,
// added to the recovered initial model (createVelDB, main-enfwi-damp.cpp:182-209).
#include <cmath>
#include <cstdio>
#include <fstream>
#include <vector>

#include "../../include/swfwi_hip.h"
#include "rsfio.h"

namespace {

void check(int rc, const char* what) {
  if (rc != SWFWI_OK) rsf::die(std::string(what) + ": " + swfwi_last_error());
}

float vel_recover(float v, float dx, float dt) { return std::sqrt(dx * dx / (dt * dt * v)); }
float vel_trans(float v, float dx, float dt) { return dx * dx / (dt * dt * v * v); }

// slownessL1L2Norm (main-enfwi-damp.cpp:368-402), a log line only: relative
// slowness misfit of a member against the real model, as written -- the
// padded models indexed with the unpadded nz over ix in [nb, nx-nb),
// iz in [0, nz-nb), FP32 sums.
void slowness_norms(const std::vector<float>& real, const float* vel, int nx, int nz, int nb, float dx, float dt,
                    float& l1, float& l2) {
  float norm1 = 0, norm2 = 0, sumSquare = 0, sum = 0;
  for (int ix = nb; ix < nx - nb; ix++)
    for (int iz = 0; iz < nz - nb; iz++) {
      const int off = ix * nz + iz;
      const float acc = 1.0f / vel_recover(real[off], dx, dt), cur = 1.0f / vel_recover(vel[off], dx, dt);
      const float d = std::abs(acc - cur);
      norm2 += d * d;
      sumSquare += acc * acc;
      norm1 += d;
      sum += acc;
    }
  l1 = norm1 / sum;
  l2 = norm2 / sumSquare;
}

}  // namespace

int main(int argc, char** argv) {
  rsf::Params par(argc, argv);
  rsf::Input vinit(par.need("vin"));
  rsf::Input vreal(par.need("vreal"));
  rsf::Input shots(par.need("shots"));
  const std::string perin = par.need("perin");
  int niter, nita, seed, nsample, niterenkf, nz, nx, nt, ng, ns, sxbeg, szbeg, gxbeg, gzbeg, jsx, jsz, jgx, jgz, nb;
  float maxdv, sigfac, dz, dx, dt, amp, fm, vmin, vmax;
  if (!par.getint("niter", niter)) rsf::die("no niter");
  if (!par.getfloat("maxdv", maxdv)) rsf::die("no maxdv");
  if (!par.getint("nita", nita)) rsf::die("no nita");
  if (!par.getint("nsample", nsample)) rsf::die("no nsample");
  if (!par.getint("niterenkf", niterenkf)) rsf::die("no niterenkf");
  if (!par.getint("seed", seed)) seed = 10;
  if (!par.getfloat("sigfac", sigfac)) rsf::die("no sigfac");
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
  if (niterenkf < 1) rsf::die("niterenkf must be >= 1");
  // Params::check (main-enfwi-damp.cpp:169-179)
  if (!(sxbeg >= 0 && szbeg >= 0 && sxbeg + (ns - 1) * jsx < nx && szbeg + (ns - 1) * jsz < nz))
    rsf::die("sources exceeds the computing zone!");
  if (!(gxbeg >= 0 && gzbeg >= 0 && gxbeg + (ng - 1) * jgx < nx && gzbeg + (ng - 1) * jgz < nz))
    rsf::die("geophones exceeds the computing zone!");

  rsf::Output vout(par.need("vout"));
  rsf::Output absobjs(par.need("absobjs"));
  rsf::Output norobjs(par.need("norobjs"));
  vout.putint("n1", nz);
  vout.putint("n2", nx);
  vout.putint("n3", niter);
  vout.putfloat("d1", dz);
  vout.putfloat("d2", dx);
  vout.putint("d3", 1);
  vout.putstring("label1", "Depth");
  vout.putstring("label2", "Distance");
  vout.putstring("label3", "Iteration");
  for (rsf::Output* o : {&absobjs, &norobjs}) {
    o->putint("n1", niter + 1);
    o->putfloat("d1", 1);
    o->putfloat("o1", 1);
  }
  absobjs.putstring("label1", "Absolute");
  norobjs.putstring("label1", "Normalize");

  // geometry (ShotPosition: sources/receivers on a regular line)
  std::vector<int> sx(ns), sz(ns), gx(ng), gz(ng);
  for (int i = 0; i < ns; i++) {
    sx[i] = sxbeg + i * jsx;
    sz[i] = szbeg + i * jsz;
  }
  for (int i = 0; i < ng; i++) {
    gx[i] = gxbeg + i * jgx;
    gz[i] = gzbeg + i * jgz;
  }
  swfwi_solver s;
  check(swfwi_solver_create(&s, nx, nz, nb, dt, dx, fm, nt, ns, sx.data(), sz.data(), ng, gx.data(), gz.data(),
                            SWFWI_PREC_EXACT),
        "solver");
  int dims[6];
  check(swfwi_solver_dims(s, dims), "dims");
  const int nxpad = dims[0], nzpad = dims[1];
  const size_t npad = (size_t)nxpad * nzpad;
  std::vector<float> v0 = vinit.floats((size_t)nx * nz), exvel(npad), exreal(npad);
  check(swfwi_expand_domain(v0.data(), nx, nz, nb, dx, dt, exvel.data()), "expandDomain");
  std::vector<float> vr = vreal.floats((size_t)nx * nz);
  check(swfwi_expand_domain(vr.data(), nx, nz, nb, dx, dt, exreal.data()), "expandDomain");
  std::vector<float> wlt(nt);
  check(swfwi_ricker(wlt.data(), nt, fm, dt, amp), "rickerWavelet");

  // ShotDataReader::serialRead: [ns][ng][nt] on disk -> [ns][nt][ng]
  std::vector<float> disk = shots.floats((size_t)ns * ng * nt);
  std::vector<float> dobs((size_t)ns * nt * ng);
  for (int is = 0; is < ns; is++)
    for (int ig = 0; ig < ng; ig++)
      for (int it = 0; it < nt; it++)
        dobs[((size_t)is * nt + it) * ng + ig] = disk[((size_t)is * ng + ig) * nt + it];

  // createVelDB: velTrans(velRecover(exvel) + perturbation) per member
  std::ifstream pf(perin, std::ios::binary);
  if (!pf) rsf::die("cannot open file: " + perin);
  std::vector<float> members((size_t)nsample * npad), tmp(npad);
  for (int iv = 0; iv < nsample; iv++) {
    pf.read(reinterpret_cast<char*>(tmp.data()), npad * sizeof(float));
    if (!pf) rsf::die("perturbation file too short: " + perin);
    for (size_t i = 0; i < npad; i++)
      members[(size_t)iv * npad + i] = vel_trans(vel_recover(exvel[i], dx, dt) + tmp[i], dx, dt);
  }

  swfwi_enkf enkf;
  check(swfwi_enkf_create(&enkf, s, wlt.data(), dobs.data(), sigfac, nsample), "EnkfAnalyze");
  check(swfwi_enkf_set_members(enkf, members.data()), "members");
  swfwi_enfwi fwi;
  check(swfwi_enfwi_create(&fwi, s, enkf, vmin, vmax, nita, maxdv), "EssFwiFramework");

  unsigned long long analyses = 0;
  std::vector<float> absobj, norobj, mean(npad), phys((size_t)nx * nz);
  auto write_mean = [&](bool write) {
    float obj = 0;
    check(swfwi_enfwi_mean_objective(fwi, mean.data(), &obj), "calobj");
    if (write) {  // sfWriteVel: physical part, recovered to m/s
      for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++)
          phys[(size_t)ix * nz + iz] = vel_recover(mean[(size_t)(ix + nb + 6) * nzpad + iz + 6], dx, dt);
      vout.floatwrite(phys.data(), phys.size());
    }
    absobj.push_back(obj);
    norobj.push_back(obj / absobj[0]);
    return obj;
  };

  check(swfwi_enkf_analyze(enkf, nullptr, (unsigned long long)seed * 1000003ull + analyses++), "pAnalyze");
  std::printf("initial analysis: mean-model objective %e\n", write_mean(false));
  std::vector<float> st((size_t)nsample * 10);
  for (int iter = 0; iter < niter; iter++) {
    check(swfwi_enfwi_epoch(fwi, iter), "epoch");
    if (iter % niterenkf == 0) {
      check(swfwi_enfwi_stats(fwi, st.data(), nullptr, nullptr), "stats");
      check(swfwi_enkf_get_members(enkf, members.data()), "members");
      for (int iv = 0; iv < nsample; iv++) {
        float l1, l2;
        slowness_norms(exreal, &members[(size_t)iv * npad], nx, nz, nb, dx, dt, l1, l2);
        std::printf("%4d/%d iter, velset[%d/%d] data rss: %g slowness l1norm: %g, l2norm: %g\n", iter, niter, iv + 1,
                    nsample, st[(size_t)iv * 10 + 6], l1, l2);
      }
      check(swfwi_enkf_analyze(enkf, nullptr, (unsigned long long)seed * 1000003ull + analyses++), "pAnalyze");
    }
    const float obj = write_mean(true);
    std::printf("iter %d, objval for mean model %e\n", iter, obj);
  }
  absobjs.floatwrite(absobj.data(), absobj.size());
  norobjs.floatwrite(norobj.data(), norobj.size());
  swfwi_enfwi_destroy(fwi);
  swfwi_enkf_destroy(enkf);
  swfwi_solver_destroy(s);
  return 0;
}
