// Device kernels of the ensemble Kalman analysis (EnkfAnalyze::pCalGainMatrix
// and pAnalyze, enkfanalyze.cpp:113-183,343-495,578-597).
//
// Layout: every data-space matrix is stored member-major, X[i*nd + r] for
// member i and data row r (r = it*ng + ig; the reference stores the transposed
// trace order, which only permutes rows and leaves every product unchanged).
#include <cfloat>
#include <cstdint>

#include "device.h"
#include "enkf.h"

namespace swfwi {
namespace {

constexpr int kB = 256;

inline int nblocks(long long n, int cap = 8192) {
  long long b = (n + kB - 1) / kB;
  if (b < 1) b = 1;
  return (int)(b > cap ? cap : b);
}

// splitmix64 finaliser: a counter-based 64-bit hash
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// N(0,1) from two hashed uniforms (Box-Muller, FP64)
__device__ __forceinline__ double normal_at(uint64_t seed, uint64_t ctr) {
  const uint64_t a = mix64(seed ^ mix64(2 * ctr));
  const uint64_t b = mix64(seed ^ mix64(2 * ctr + 1));
  const double u1 = ((a >> 11) + 1.0) * (1.0 / 9007199254740993.0);  // (0, 1]
  const double u2 = (b >> 11) * (1.0 / 9007199254740992.0);          // [0, 1)
  return sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}

// HA' = HA - rowmean(HA) in FP64 (pInitGamma on HOnA, enkfanalyze.cpp:520-545),
// plus per-block arg-max of |HA'| compared as float (abs_less<float>) with
// the first index winning ties (std::max_element).
__global__ void k_ha_perturb(const float* ha, int N, long long nd, double* hap, float* blk_abs, long long* blk_idx) {
  __shared__ float sa[kB];
  __shared__ long long si[kB];
  float best = -1.f;
  long long bidx = -1;
  for (long long r = blockIdx.x * (long long)kB + threadIdx.x; r < nd; r += (long long)gridDim.x * kB) {
    double s = 0.0;
    for (int i = 0; i < N; i++) s += ha[(long long)i * nd + r];
    const double avg = s / N;
    for (int i = 0; i < N; i++) {
      const double v = ha[(long long)i * nd + r] - avg;
      hap[(long long)i * nd + r] = v;
      const float av = fabsf((float)v);
      const long long idx = (long long)i * nd + r;
      if (av > best || (av == best && idx < bidx)) {
        best = av;
        bidx = idx;
      }
    }
  }
  sa[threadIdx.x] = best;
  si[threadIdx.x] = bidx;
  __syncthreads();
  for (int o = kB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      const float a = sa[threadIdx.x + o];
      const long long ia = si[threadIdx.x + o];
      if (a > sa[threadIdx.x] || (a == sa[threadIdx.x] && ia >= 0 && (si[threadIdx.x] < 0 || ia < si[threadIdx.x]))) {
        sa[threadIdx.x] = a;
        si[threadIdx.x] = ia;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    blk_abs[blockIdx.x] = sa[0];
    blk_idx[blockIdx.x] = si[0];
  }
}

// Perturbation E (generated N(0, sigma) or given), D = d + E, gamma = E -
// rowmean(E), band = HA' + gamma, t0 = D - HA (enkfanalyze.cpp:430-460).
__global__ void k_band(const float* ha, const double* hap, const float* d, const double* e_in, uint64_t seed,
                       double sigma, int N, long long nd, double* band, double* t0) {
  for (long long r = blockIdx.x * (long long)kB + threadIdx.x; r < nd; r += (long long)gridDim.x * kB) {
    double s = 0.0;
    for (int i = 0; i < N; i++) {
      const double e = e_in ? e_in[(long long)i * nd + r] : sigma * normal_at(seed, (uint64_t)i * nd + r);
      s += e;
      band[(long long)i * nd + r] = e;  // E parked in band until its row mean is known
    }
    const double avg = s / N;
    const double dr = d[r];
    for (int i = 0; i < N; i++) {
      const long long q = (long long)i * nd + r;
      const double e = band[q];
      t0[q] = (dr + e) - (double)ha[q];
      band[q] = hap[q] + (e - avg);
    }
  }
}

// Per point of the model: pCreateAMean (FP32 member-order sum / N), A' = A -
// mean in FP64, t5 = A' * T, then vel = velTrans(velRecover(vel) + t5) in the
// reference's float/double mix (enkfanalyze.cpp:113-183,578-597, common.h:21-30).
constexpr int kChunk = 16;  // members updated per pass (register accumulators)

__global__ void k_model_update(const float* vels, float* out, long long n, int N, const double* __restrict__ sT,
                               float dx, float dt) {
  // sT (N x N) is read with wave-uniform indices: served from the scalar cache
  for (long long p = blockIdx.x * (long long)kB + threadIdx.x; p < n; p += (long long)gridDim.x * kB) {
    float sum = 0.f;
    for (int j = 0; j < N; j++) sum += vels[(long long)j * n + p];
    const float mean = sum / N;
    for (int i0 = 0; i0 < N; i0 += kChunk) {
      double acc[kChunk];
#pragma unroll
      for (int k = 0; k < kChunk; k++) acc[k] = 0.0;
      for (int j = 0; j < N; j++) {
        const double ap = (double)vels[(long long)j * n + p] - (double)mean;
#pragma unroll
        for (int k = 0; k < kChunk; k++)
          if (i0 + k < N) acc[k] = __builtin_fma(ap, sT[j + N * (i0 + k)], acc[k]);
      }
#pragma unroll
      for (int k = 0; k < kChunk; k++) {
        const int i = i0 + k;
        if (i >= N) break;
        float v = vels[(long long)i * n + p];
        float t = dx * dx / (dt * dt * v);  // velRecover
        v = sqrtf(t);
        v = (float)((double)v + acc[k]);
        v = dx * dx / (dt * dt * v * v);    // velTrans
        out[(long long)i * n + p] = v;
      }
    }
  }
}

// FP32 member-order mean (pCreateAMean)
__global__ void k_member_mean(const float* vels, long long n, int N, float* mean) {
  for (long long p = blockIdx.x * (long long)kB + threadIdx.x; p < n; p += (long long)gridDim.x * kB) {
    float sum = 0.f;
    for (int j = 0; j < N; j++) sum += vels[(long long)j * n + p];
    mean[p] = sum / N;
  }
}

// float -> double copy (member-major data matrices)
__global__ void k_to_double(const float* a, double* b, long long n) {
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB) b[i] = a[i];
}

}  // namespace

void enkf_ha_perturb(const float* ha, int N, long long nd, double* hap, double* max_abs_value, hipStream_t s) {
  const int nb = nblocks(nd, 1024);
  DBuf<float> ba((size_t)nb);
  DBuf<long long> bi((size_t)nb);
  hipLaunchKernelGGL(k_ha_perturb, dim3(nb), dim3(kB), 0, s, ha, N, nd, hap, ba.get(), bi.get());
  std::vector<float> ha_abs(nb);
  std::vector<long long> ha_idx(nb);
  ba.download(ha_abs.data(), nb, s);
  bi.download(ha_idx.data(), nb, s);
  SWFWI_HIP(hipStreamSynchronize(s));
  float best = -1.f;
  long long idx = -1;
  for (int b = 0; b < nb; b++) {
    if (ha_idx[b] < 0) continue;
    if (ha_abs[b] > best || (ha_abs[b] == best && ha_idx[b] < idx)) {
      best = ha_abs[b];
      idx = ha_idx[b];
    }
  }
  double v = 0;
  if (idx >= 0) SWFWI_HIP(hipMemcpy(&v, hap + idx, sizeof(double), hipMemcpyDeviceToHost));
  *max_abs_value = v < 0 ? -v : v;
}

void enkf_band(const float* ha, const double* hap, const float* d, const double* e_in, uint64_t seed, double sigma,
               int N, long long nd, double* band, double* t0, hipStream_t s) {
  hipLaunchKernelGGL(k_band, dim3(nblocks(nd)), dim3(kB), 0, s, ha, hap, d, e_in, seed, sigma, N, nd, band, t0);
}

void enkf_model_update(const float* vels, float* out, long long n, int N, const double* T, float dx, float dt,
                       hipStream_t s) {
  if (N > kMaxMembers) fail(kErrRange, "ensemble larger than the supported maximum");
  hipLaunchKernelGGL(k_model_update, dim3(nblocks(n, 2048)), dim3(kB), 0, s, vels, out, n, N, T, dx, dt);
}

void enkf_member_mean(const float* vels, long long n, int N, float* mean, hipStream_t s) {
  hipLaunchKernelGGL(k_member_mean, dim3(nblocks(n)), dim3(kB), 0, s, vels, n, N, mean);
}

void enkf_to_double(const float* a, double* b, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_to_double, dim3(nblocks(n)), dim3(kB), 0, s, a, b, n);
}

}  // namespace swfwi
