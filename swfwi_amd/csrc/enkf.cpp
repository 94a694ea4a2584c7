#include "enkf.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace swfwi {

bool jacobi_eigen(std::vector<double>& a, int n, std::vector<double>& w, std::vector<double>& v) {
  // Cyclic Jacobi rotations (Golub & Van Loan, Alg. 8.5.3) until the
  // off-diagonal Frobenius norm is negligible against the whole matrix.
  auto A = [&](int i, int j) -> double& { return a[(size_t)i + (size_t)n * j]; };
  v.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) v[(size_t)i + (size_t)n * i] = 1.0;
  double total = 0;
  for (double x : a) total += x * x;
  bool converged = false;
  for (int sweep = 0; sweep < 100 && !converged; sweep++) {
    double off = 0;
    for (int j = 0; j < n; j++)
      for (int i = 0; i < j; i++) off += 2 * A(i, j) * A(i, j);
    if (off <= 1e-30 * total || off == 0) {
      converged = true;
      break;
    }
    for (int p = 0; p < n - 1; p++) {
      for (int q = p + 1; q < n; q++) {
        const double apq = A(p, q);
        if (apq == 0.0) continue;
        const double app = A(p, p), aqq = A(q, q);
        const double theta = (aqq - app) / (2 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        const double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; k++) {  // A <- A J (columns p, q)
          const double akp = A(k, p), akq = A(k, q);
          A(k, p) = c * akp - s * akq;
          A(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {  // A <- J^T A (rows p, q)
          const double apk = A(p, k), aqk = A(q, k);
          A(p, k) = c * apk - s * aqk;
          A(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          double& vkp = v[(size_t)k + (size_t)n * p];
          double& vkq = v[(size_t)k + (size_t)n * q];
          const double x = vkp, y = vkq;
          vkp = c * x - s * y;
          vkq = s * x + c * y;
        }
      }
    }
  }
  // sort descending
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int x, int y) { return A(x, x) > A(y, y); });
  std::vector<double> vs((size_t)n * n);
  w.resize(n);
  for (int k = 0; k < n; k++) {
    w[k] = A(order[k], order[k]);
    for (int i = 0; i < n; i++) vs[(size_t)i + (size_t)n * k] = v[(size_t)i + (size_t)n * order[k]];
  }
  v.swap(vs);
  return converged;
}

int clip_position(const std::vector<double>& s) {
  // Matrix.cpp:170-185
  const double relativeError = 0.001f;
  const int N = (int)s.size();
  double sum = 0;
  int pos;
  for (pos = 0; pos < N - 1; pos++) {
    sum += std::abs(s[pos]);
    if (relativeError / (1 - relativeError) * sum > (N - pos + 1) * s[pos + 1]) break;
  }
  return pos + 1;
}

Enkf::Enkf(Solver& s, const float* wlt_host, const float* dobs_host, float sigfac, int nmembers)
    : s_(s), N_(nmembers), sigfac_(sigfac) {
  if (nmembers < 2 || nmembers > kMaxMembers) fail(kErrRange, "ensemble size must be in [2, 1024]");
  const Geometry& g = s.geo();
  s.set_wavelet(wlt_host);
  dobs_.upload(dobs_host, (size_t)g.ns() * g.nt * g.ng(), s.stream());
  mem_.resize((size_t)N_ * s.npts());
  mem_out_.resize((size_t)N_ * s.npts());
  s.prop().sync();
}

void Enkf::set_members(const float* host) {
  mem_.upload(host, (size_t)N_ * s_.npts(), s_.stream());
  s_.prop().sync();
}

void Enkf::get_members(float* host) {
  mem_.download(host, (size_t)N_ * s_.npts(), s_.stream());
  s_.prop().sync();
}

void Enkf::mean(float* host) {
  mean_.resize((size_t)s_.npts());
  enkf_member_mean(mem_.get(), s_.npts(), N_, mean_.get(), s_.stream());
  mean_.download(host, (size_t)s_.npts(), s_.stream());
  s_.prop().sync();
}

void Enkf::analyze(const double* e_host, uint64_t seed) {
  const Geometry& g = s_.geo();
  hipStream_t st = s_.stream();
  const int ns = g.ns(), nt = g.nt, ng = g.ng(), N = N_;
  const long long n = s_.npts(), nd = (long long)nt * ng;

  // codes of this analysis (pAnalyze draws one set, enkfanalyze.cpp:113-115)
  const std::vector<int> code = codes_.next(ns);
  code_.upload(code, st);
  encsrc_.resize((size_t)nt * ns);
  d_.resize((size_t)nd);
  op_encode_source(s_.wavelet(), code_.get(), ns, nt, encsrc_.get(), st);
  op_encode_obs(dobs_.get(), code_.get(), ns, nd, d_.get(), st);

  // H on every member: one batched encoded forward run, no mute
  // (pCalGainMatrix, enkfanalyze.cpp:367-398)
  rv_.resize((size_t)N * n);
  op_reciprocal(mem_.get(), rv_.get(), (long long)N * n, st);
  ha_.resize((size_t)N * nd);
  Propagator::Sources src{encsrc_.get(), ns, 0, nullptr, nullptr};
  s_.prop().forward(N, rv_.get(), n, src, ha_.get(), nullptr, w_);

  // HA' and the perturbation scale fixed at the first analysis
  hap_.resize((size_t)N * nd);
  double maxabs = 0;
  enkf_ha_perturb(ha_.get(), N, nd, hap_.get(), &maxabs, st);
  if (!sigma_init_) {
    sigma_init_ = true;
    sigma_ = maxabs * sigfac_;  // initPerturbSigma (enkfanalyze.cpp:497-499)
  }
  const double* e_dev = nullptr;
  if (e_host) {
    e_.upload(e_host, (size_t)N * nd, st);
    e_dev = e_.get();
  }
  band_.resize((size_t)N * nd);
  t0_.resize((size_t)N * nd);
  enkf_band(ha_.get(), hap_.get(), d_.get(), e_dev, seed, sigma_, N, nd, band_.get(), t0_.get(), st);

  // the three N x N FP64 data-space products
  G_.resize((size_t)N * N);
  P_.resize((size_t)N * N);
  Q_.resize((size_t)N * N);
  op_gram_f64(band_.get(), band_.get(), N, nd, G_.get(), st);
  op_gram_f64(hap_.get(), band_.get(), N, nd, P_.get(), st);
  op_gram_f64(band_.get(), t0_.get(), N, nd, Q_.get(), st);
  std::vector<double> G((size_t)N * N), P((size_t)N * N), Q((size_t)N * N);
  G_.download(G.data(), G.size(), st);
  P_.download(P.data(), P.size(), st);
  Q_.download(Q.data(), Q.size(), st);
  SWFWI_HIP(hipStreamSynchronize(st));

  // band = U S V^T  <=>  band^T band = V S^2 V^T; then
  // t4 = HA'^T U_k S_k^-2 U_k^T (D - HA) = P V_k S_k^-4 V_k^T Q
  std::vector<double> lam, V;
  if (!jacobi_eigen(G, N, lam, V)) fail(kErrNumeric, "The algorithm computing SVD failed to converge.");
  svals_.resize(N);
  for (int k = 0; k < N; k++) svals_[k] = std::sqrt(std::max(lam[k], 0.0));
  for (double sv : svals_)
    if (!std::isfinite(sv)) fail(kErrNumeric, "The algorithm computing SVD failed to converge.");
  clip_ = clip_position(svals_);
  std::vector<double> PV((size_t)N * clip_), VQ((size_t)clip_ * N);
  for (int k = 0; k < clip_; k++) {
    const double inv = (1 / svals_[k]) * (1 / svals_[k]);
    const double scale = inv * inv;
    for (int j = 0; j < N; j++) {
      double a = 0;
      for (int m = 0; m < N; m++) a += P[(size_t)j + (size_t)N * m] * V[(size_t)m + (size_t)N * k];
      PV[(size_t)j + (size_t)N * k] = a * scale;
    }
    for (int i = 0; i < N; i++) {
      double b = 0;
      for (int m = 0; m < N; m++) b += V[(size_t)m + (size_t)N * k] * Q[(size_t)m + (size_t)N * i];
      VQ[(size_t)k + (size_t)clip_ * i] = b;
    }
  }
  t4_.assign((size_t)N * N, 0.0);
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      double acc = 0;
      for (int k = 0; k < clip_; k++) acc += PV[(size_t)j + (size_t)N * k] * VQ[(size_t)k + (size_t)clip_ * i];
      t4_[(size_t)j + (size_t)N * i] = acc;
    }

  // model space: A' t4 and the velTrans(velRecover(v) + t5) update
  T_.upload(t4_, st);
  const Geometry& gg = s_.geo();
  enkf_model_update(mem_.get(), mem_out_.get(), n, N, T_.get(), gg.dx, gg.dt, st);
  std::swap(mem_, mem_out_);
  SWFWI_HIP(hipStreamSynchronize(st));
}

}  // namespace swfwi
