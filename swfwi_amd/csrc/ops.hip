// Data-side kernels of the FWI drivers.  They are all HBM-streaming,
// elementwise or reduction kernels: coalesced grid-stride loops, FP64
// accumulation through wavefront shuffles, fixed reduction order.
#include <cfloat>

#include "device.h"
#include "ops.h"

namespace swfwi {
namespace {

constexpr int kB = 256;

inline int nblocks(long long n, int cap = 4096) {
  long long b = (n + kB - 1) / kB;
  if (b < 1) b = 1;
  return (int)(b > cap ? cap : b);
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum over a 256-thread block; the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0) r = ((sh[0] + sh[1]) + sh[2]) + sh[3];
  __syncthreads();
  return r;
}

__global__ void k_encode_obs(const float* dobs, const int* code, int ns, long long m, float* out) {
  for (long long j = blockIdx.x * (long long)kB + threadIdx.x; j < m; j += (long long)gridDim.x * kB) {
    float acc = 0;
    for (int s = 0; s < ns; s++) acc = acc + dobs[(long long)s * m + j] * (float)code[s];
    out[j] = acc;
  }
}

__global__ void k_encode_source(const float* wlt, const int* code, int ns, int nt, float* out) {
  long long n = (long long)nt * ns;
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB) {
    int it = (int)(i / ns), s = (int)(i - (long long)it * ns);
    out[i] = wlt[it] * (float)code[s];
  }
}

constexpr int kMuteChunk = 4096;

__global__ void k_mute_windows(const float* vel, long long vstride, GridGeom g, const int* sxp, const int* szp,
                               const int* gxp, const int* gzp, int ng, int nt, int half_len, int* start, int* end) {
  __shared__ float buf[kMuteChunk];
  __shared__ float avg_sh;
  const int e = blockIdx.x;
  const float* v = vel + (long long)e * vstride;
  const int sx = sxp[e], sz = szp[e], gz0 = gzp[0];
  const int gmin = sz < gz0 ? sz : gz0, gmax = sz > gz0 ? sz : gz0;
  const int rows = gmax - gmin + 1;
  const long long cnt = (long long)g.nx * rows;
  float acc = 0.0f;  // meaningful in thread 0: the reference's serial FP32 order
  for (long long base = 0; base < cnt; base += kMuteChunk) {
    const int len = (int)((cnt - base) < kMuteChunk ? (cnt - base) : kMuteChunk);
    for (int j = threadIdx.x; j < len; j += kB) {
      const long long idx = base + j;
      const int i = (int)(idx / rows), k = gmin + (int)(idx - (long long)i * rows);
      buf[j] = v[(long long)i * g.ld + k];
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int j = 0; j < len; j++) acc += buf[j];
    __syncthreads();
  }
  if (threadIdx.x == 0) avg_sh = acc / (g.nx * rows);
  __syncthreads();
  const float avg = avg_sh;
  for (int r = threadIdx.x; r < ng; r += kB) {
    const int gx = gxp[r], gz = gzp[r];
    const float dist = (gx - sx) * (gx - sx) + (gz - sz) * (gz - sz);
    const int t = (int)sqrtf(dist * avg);
    start[(long long)e * ng + r] = t;
    end[(long long)e * ng + r] = ((t + 2 * half_len) > nt) ? nt : (t + 2 * half_len);
  }
}

__global__ void k_mute_apply(float* data, long long dstride, int E, const int* start, const int* end, long long wstride,
                             int nt, int ng) {
  const long long per = (long long)nt * ng, n = per * E;
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB) {
    const int e = (int)(i / per);
    const long long j = i - (long long)e * per;
    const int t = (int)(j / ng), r = (int)(j - (long long)t * ng);
    const long long w = (long long)e * wstride + r;
    if (t >= start[w] && t < end[w]) data[(long long)e * dstride + j] = 0.f;
  }
}

__global__ void k_residual(const float* obs, long long ostride, const float* cal, float* res, long long m,
                           double* partials) {
  __shared__ double sh[4];
  const int e = blockIdx.y;
  const float* o = obs + (long long)e * ostride;
  const float* c = cal + (long long)e * m;
  float* r = res + (long long)e * m;
  double acc = 0;
  for (long long j = blockIdx.x * (long long)kB + threadIdx.x; j < m; j += (long long)gridDim.x * kB) {
    const float d = o[j] - c[j];
    r[j] = d;
    acc += (double)d * (double)d;
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[(long long)e * gridDim.x + blockIdx.x] = acc;
}

__global__ void k_sumsq(const float* x, long long m, double* partials) {
  __shared__ double sh[4];
  const int e = blockIdx.y;
  const float* p = x + (long long)e * m;
  double acc = 0;
  for (long long j = blockIdx.x * (long long)kB + threadIdx.x; j < m; j += (long long)gridDim.x * kB)
    acc += (double)p[j] * (double)p[j];
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[(long long)e * gridDim.x + blockIdx.x] = acc;
}

__global__ void k_reduce_partials(const double* partials, int nblk, double* out) {
  __shared__ double sh[4];
  const int e = blockIdx.x;
  double acc = 0;
  for (int j = threadIdx.x; j < nblk; j += kB) acc += partials[(long long)e * nblk + j];
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) out[e] = acc;
}

__global__ void k_trans_vsrc(const float* in, float* out, int nt, int ng, int E) {
  const long long per = (long long)nt * ng, n = per * E;
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB) {
    const int e = (int)(i / per);
    const long long j = i - (long long)e * per;
    const int t = (int)(j / ng);
    float v;
    if (t <= 1 || t == nt - 1 || t == nt - 2) {
      v = 0.0f;
    } else {
      const float* p = in + i;
      v = -1. / 12 * p[-2LL * ng] + 4. / 3 * p[-1LL * ng] - 2.5 * p[0] + 4. / 3 * p[ng] - 1. / 12 * p[2LL * ng];
    }
    out[i] = v;
  }
}

__global__ void k_scale_mask(float* grad, long long gstride, int E, GridGeom g) {
  const long long per = (long long)g.nx * g.ld, n = per * E;
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB) {
    const int e = (int)(i / per);
    const long long j = i - (long long)e * per;
    const int ix = (int)(j / g.ld), iz = (int)(j - (long long)ix * g.ld);
    if (iz >= g.nz) continue;
    float* p = grad + (long long)e * gstride + j;
    const bool keep = ix >= g.bx0 && iz >= g.bz0 && ix < g.nx - g.bxn && iz < g.nz - g.bzn;
    if (!keep) {
      *p = 0.f;
    } else {
      const int k = iz - g.bz0;
      if (k >= 1) *p = *p * sqrtf((float)k);
    }
  }
}

__global__ void k_sum_entries(const float* in, long long stride, int E, long long n, float* out) {
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB) {
    float acc = 0;
    for (int e = 0; e < E; e++) acc = acc + in[(long long)e * stride + i];
    out[i] = acc;
  }
}

__global__ void k_cg_dots(const float* cur, const float* pre, long long n, double* partials) {
  __shared__ double sh[4];
  double a = 0, b = 0, c = 0;
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB) {
    const double x = cur[i], y = pre[i];
    a += x * x;
    b += x * y;
    c += y * y;
  }
  a = block_sum(a, sh);
  b = block_sum(b, sh);
  c = block_sum(c, sh);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = a;
    partials[gridDim.x + blockIdx.x] = b;
    partials[2 * gridDim.x + blockIdx.x] = c;
  }
}

__global__ void k_cg_update(const float* cur, float* pre, float* dir, float beta, long long n) {
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB) {
    const float c = cur[i];
    dir[i] = c + beta * dir[i];
    pre[i] = c;
  }
}

__device__ __forceinline__ float min_keep(float m, float x) { return (m > x) ? x : m; }

__global__ void k_max_alpha(const float* vel, const float* grad, long long n, float dt, float dx, float maxdv,
                            float* partial) {
  __shared__ float sh[kB];
  float m = FLT_MAX;
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB) {
    const float v = vel[i];
    float tmpv = dx / (dt * sqrtf(v));
    tmpv -= maxdv;
    tmpv = (dx / (dt * tmpv)) * (dx / (dt * tmpv));
    const float ag = fabsf(grad[i]);
    if ((double)ag < 1e-10) continue;
    m = min_keep(m, (tmpv - v) / ag);
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int o = kB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] = min_keep(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void k_min_partials(const float* partial, int nblk, float* out) {
  __shared__ float sh[kB];
  float m = FLT_MAX;
  for (int j = threadIdx.x; j < nblk; j += kB) m = min_keep(m, partial[j]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int o = kB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] = min_keep(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

struct Steps {
  float v[16];
};

__global__ void k_update_vel(const float* v, const float* dir, Steps st, int E, long long n, float tmin, float tmax,
                             float* nv, long long nvstride) {
  const long long tot = n * E;
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < tot; i += (long long)gridDim.x * kB) {
    const int e = (int)(i / n);
    const long long j = i - (long long)e * n;
    float x = v[j] + st.v[e] * dir[j];
    if (x > tmax) x = tmax;
    if (x < tmin) x = tmin;
    nv[(long long)e * nvstride + j] = x;
  }
}

__global__ void k_refill(float* a, long long stride, int E, GridGeom g) {
  const long long per = (long long)g.nx * g.ld, n = per * E;
  const int xr = g.nx - g.bxn - 1, zr = g.nz - g.bzn - 1;
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB) {
    const int e = (int)(i / per);
    const long long j = i - (long long)e * per;
    const int ix = (int)(j / g.ld), iz = (int)(j - (long long)ix * g.ld);
    if (iz >= g.nz) continue;
    const int cx = ix < g.bx0 ? g.bx0 : (ix > xr ? xr : ix);
    const int cz = iz < g.bz0 ? g.bz0 : (iz > zr ? zr : iz);
    if (cx != ix || cz != iz) {
      float* b = a + (long long)e * stride;
      b[j] = b[(long long)cx * g.ld + cz];
    }
  }
}

// The 6-wide boundary band of writeBndry/readBndry (damp4t10d.cpp:478-536):
// band offset i -> grid point, bottom rows first (iz fastest), then the left
// and the right columns (iz fastest).
__device__ __forceinline__ long long band_point(const GridGeom& g, int i) {
  const int w = 6;
  const int nxb = g.nx - (g.bx0 - w + g.bxn - w), nzb = g.nz - g.bz0 - g.bzn;
  int ix, iz;
  if (i < w * nxb) {
    ix = g.bx0 - w + i / w;
    iz = g.nz - g.bzn + i % w;
  } else {
    const int k = i - w * nxb;
    const int col = k / nzb;
    iz = g.bz0 + k % nzb;
    ix = col < w ? g.bx0 - w + col : g.nx - g.bxn + (col - w);
  }
  return (long long)ix * g.ld + iz;
}
__global__ void k_band_pack(const float* p, float* band, int bsize, GridGeom g) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i < bsize) band[i] = p[band_point(g, i)];
}
__global__ void k_band_unpack(const float* band, float* p, int bsize, GridGeom g) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i < bsize) p[band_point(g, i)] = band[i];
}

__global__ void k_reciprocal2(const float* vel, float* rv, long long n) {
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB)
    rv[i] = 1.0f / vel[i];
}

__global__ void k_fill_any(float* p, long long n) {
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB)
    p[i] = (float)(i & 1023);
}

__global__ void k_axpy1(const float* src, float* dst, long long n) {
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB)
    dst[i] = dst[i] + src[i];
}

// C (N x N, column-major) += A_i . B_j over m, tiled 16x16 outputs per block,
// rows in chunks staged through LDS; partial products go to part[chunk][N*N]
// and are summed in chunk order by k_gram_finish.
constexpr int kGT = 16;
constexpr int kGK = 64;

__global__ void k_gram(const double* A, const double* B, int N, long long m, long long chunk, double* part) {
  __shared__ double sa[kGT][kGK + 1];
  __shared__ double sb[kGT][kGK + 1];
  const int ti = threadIdx.x % kGT, tj = threadIdx.x / kGT;
  const int i0 = blockIdx.x * kGT, j0 = blockIdx.y * kGT;
  const long long r0 = (long long)blockIdx.z * chunk;
  const long long r1 = r0 + chunk < m ? r0 + chunk : m;
  double acc = 0;
  for (long long rb = r0; rb < r1; rb += kGK) {
    for (int t = threadIdx.x; t < kGT * kGK; t += kGT * kGT) {
      const int a = t / kGK, k = t - a * kGK;
      const long long r = rb + k;
      sa[a][k] = (i0 + a < N && r < r1) ? A[(long long)(i0 + a) * m + r] : 0.0;
      sb[a][k] = (j0 + a < N && r < r1) ? B[(long long)(j0 + a) * m + r] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int k = 0; k < kGK; k++) acc = __builtin_fma(sa[ti][k], sb[tj][k], acc);
    __syncthreads();
  }
  const int i = i0 + ti, j = j0 + tj;
  if (i < N && j < N) part[(long long)blockIdx.z * N * N + i + (long long)N * j] = acc;
}

// FP64 MFMA forms for ensembles of up to kGMaxN members.
// v_mfma_f64_16x16x4f64: lane l supplies A[i = l%16][k = l/16] and
// B[k = l/16][j = l%16] and holds D[i = l/16 + 4*v][j = l%16] in accumulator
// item v (measured).
constexpr int kGMaxN = 112;  // 7 tiles of 16 members
typedef double v4d __attribute__((ext_vector_type(4)));

// FP64 MFMA Gram product, one wave per 16-member tile row: wave I forms the
// tiles (I, J) of X^T Y for J = 0 .. T-1 (J >= I when X == Y, the upper
// triangle), so per K-step it reads its row's A operand once and one B
// operand per tile.  Each block takes a run of data rows in chunks of
// kG2KC; the next chunk is loaded into registers while the current one is
// multiplied from LDS, so the HBM latency of a chunk overlaps the products
// of the one before.  Output: block b's partial X^T Y at part[b][i + N j]
// (for X == Y only the tiles with J >= I).
constexpr int kG2KC = 16;          // data rows per chunk
constexpr int kG2Ld = kG2KC + 1;   // LDS row stride (doubles)
constexpr int kG2MaxT = kGMaxN / 16;
constexpr int kG2Threads = kG2MaxT * 64;
constexpr int kG2PerThread = kGMaxN * kG2KC / kG2Threads;  // staged values per thread and matrix
static_assert(kGMaxN * kG2KC % kG2Threads == 0, "chunk splits evenly over the block");

template <bool SYM>
__global__ __launch_bounds__(kG2Threads, 4) void k_gram_rows(const double* __restrict__ X, const double* __restrict__ Y,
                                                          int N, long long m, long long rows_per_block,
                                                          double* __restrict__ part) {
  __shared__ double sx[kGMaxN * kG2Ld];
  __shared__ double sy[SYM ? 1 : kGMaxN * kG2Ld];
  const double* sY = SYM ? sx : sy;
  const int T = (N + 15) / 16;
  const int lane = threadIdx.x & 63, I = threadIdx.x >> 6;  // this wave's tile row
  const int li = lane & 15, lk = lane >> 4;
  const long long r0 = (long long)blockIdx.x * rows_per_block;
  const long long r1 = r0 + rows_per_block < m ? r0 + rows_per_block : m;
  // staging: value c of this thread is member q / kG2KC, row q % kG2KC of
  // the chunk, q = threadIdx.x + c * kG2Threads (16 consecutive threads read
  // one member's 128 contiguous bytes)
  double nx[kG2PerThread], ny[kG2PerThread];
  auto fetch = [&](long long rb) {
#pragma unroll
    for (int c = 0; c < kG2PerThread; c++) {
      const int q = threadIdx.x + c * kG2Threads, i = q / kG2KC;
      const long long r = rb + (q % kG2KC);
      const bool in = i < N && r < r1;
      nx[c] = in ? X[(long long)i * m + r] : 0.0;
      if (!SYM) ny[c] = in ? Y[(long long)i * m + r] : 0.0;
    }
  };
  v4d acc[kG2MaxT];
#pragma unroll
  for (int J = 0; J < kG2MaxT; J++) acc[J] = v4d{0, 0, 0, 0};
  fetch(r0);
  for (long long rb = r0; rb < r1; rb += kG2KC) {
#pragma unroll
    for (int c = 0; c < kG2PerThread; c++) {
      const int q = threadIdx.x + c * kG2Threads;
      sx[(q / kG2KC) * kG2Ld + q % kG2KC] = nx[c];
      if (!SYM) sy[(q / kG2KC) * kG2Ld + q % kG2KC] = ny[c];
    }
    __syncthreads();
    if (rb + kG2KC < r1) fetch(rb + kG2KC);  // block-uniform
    if (I < T) {
      const double* pa = sx + (I * 16 + li) * kG2Ld + lk;
#pragma unroll
      for (int kk = 0; kk < kG2KC; kk += 4) {
        const double a = pa[kk];
#pragma unroll
        for (int J = 0; J < kG2MaxT; J++)
          if (J < T && (!SYM || J >= I))  // wave-uniform
            acc[J] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, sY[(J * 16 + li) * kG2Ld + lk + kk], acc[J], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  if (I >= T) return;
  double* out = part + (long long)blockIdx.x * N * N;
  const int iq = I * 16 + lk;
#pragma unroll
  for (int J = 0; J < kG2MaxT; J++) {
    if (J >= T || (SYM && J < I)) continue;
    const int j = J * 16 + li;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int i = iq + 4 * v;
      if (i < N && j < N) out[i + (long long)N * j] = acc[J][v];
    }
  }
}

// sum of the per-block partials in block order; for a symmetric product only
// the upper-triangle tiles were formed, the lower ones are mirrored
__global__ void k_gram_mfma_finish(const double* part, int nchunk, int N, bool sym, double* C) {
  const int NN = N * N;
  for (int q = blockIdx.x * kB + threadIdx.x; q < NN; q += gridDim.x * kB) {
    const int i = q % N, j = q / N;
    const int src = (sym && i / 16 > j / 16) ? j + N * i : q;
    double s = 0;
    for (int c = 0; c < nchunk; c++) s += part[(long long)c * NN + src];
    C[q] = s;
  }
}

__global__ void k_gram_finish(const double* part, int nchunk, int NN, double* C) {
  for (int q = blockIdx.x * kB + threadIdx.x; q < NN; q += gridDim.x * kB) {
    double s = 0;
    for (int c = 0; c < nchunk; c++) s += part[(long long)c * NN + q];
    C[q] = s;
  }
}

// out[r][e] ... out is (n x N) entry-major: out[e*n + r] = sum_j Ap[j*n + r] * T[j + N*e]
__global__ void k_apply_nxn(const double* Ap, const double* T, int N, long long n, double* out) {
  extern __shared__ double sT[];
  for (int q = threadIdx.x; q < N * N; q += kB) sT[q] = T[q];
  __syncthreads();
  for (long long r = blockIdx.x * (long long)kB + threadIdx.x; r < n; r += (long long)gridDim.x * kB) {
    for (int e = 0; e < N; e++) {
      double s = 0;
      for (int j = 0; j < N; j++) s = __builtin_fma(Ap[(long long)j * n + r], sT[j + N * e], s);
      out[(long long)e * n + r] = s;
    }
  }
}

}  // namespace

void op_encode_obs(const float* dobs, const int* code, int ns, long long m, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_encode_obs, dim3(nblocks(m)), dim3(kB), 0, s, dobs, code, ns, m, out);
}
void op_encode_source(const float* wlt, const int* code, int ns, int nt, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_encode_source, dim3(nblocks((long long)nt * ns)), dim3(kB), 0, s, wlt, code, ns, nt, out);
}
void op_mute_windows(const float* vel, long long vstride, int E, const GridGeom& g, const int* sxp, const int* szp,
                     const int* gxp, const int* gzp, int ng, int nt, int half_len, int* start, int* end,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_mute_windows, dim3(E), dim3(kB), 0, s, vel, vstride, g, sxp, szp, gxp, gzp, ng, nt, half_len,
                     start, end);
}
void op_mute_apply(float* data, long long dstride, int E, const int* start, const int* end, long long wstride, int nt,
                   int ng, hipStream_t s) {
  hipLaunchKernelGGL(k_mute_apply, dim3(nblocks((long long)nt * ng * E)), dim3(kB), 0, s, data, dstride, E, start,
                     end, wstride, nt, ng);
}
int op_residual(const float* obs, long long ostride, const float* cal, float* res, long long m, int E,
                double* partials, hipStream_t s) {
  const int nb = nblocks(m, 512);
  hipLaunchKernelGGL(k_residual, dim3(nb, E), dim3(kB), 0, s, obs, ostride, cal, res, m, partials);
  return nb;
}
int op_sumsq(const float* x, long long m, int E, double* partials, hipStream_t s) {
  const int nb = nblocks(m, 512);
  hipLaunchKernelGGL(k_sumsq, dim3(nb, E), dim3(kB), 0, s, x, m, partials);
  return nb;
}
void op_reduce_partials(const double* partials, int nblk, int E, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_partials, dim3(E), dim3(kB), 0, s, partials, nblk, out);
}
void op_trans_vsrc(const float* in, float* out, int nt, int ng, int E, hipStream_t s) {
  hipLaunchKernelGGL(k_trans_vsrc, dim3(nblocks((long long)nt * ng * E)), dim3(kB), 0, s, in, out, nt, ng, E);
}
void op_scale_mask(float* grad, long long gstride, int E, const GridGeom& g, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_mask, dim3(nblocks((long long)g.nx * g.ld * E)), dim3(kB), 0, s, grad, gstride, E, g);
}
void op_sum_entries(const float* in, long long stride, int E, long long n, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_entries, dim3(nblocks(n)), dim3(kB), 0, s, in, stride, E, n, out);
}
int op_cg_dots(const float* cur, const float* pre, long long n, double* partials, hipStream_t s) {
  const int nb = nblocks(n, 512);
  hipLaunchKernelGGL(k_cg_dots, dim3(nb), dim3(kB), 0, s, cur, pre, n, partials);
  return nb;
}
void op_cg_update(const float* cur, float* pre, float* dir, float beta, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_update, dim3(nblocks(n)), dim3(kB), 0, s, cur, pre, dir, beta, n);
}
void op_max_alpha(const float* vel, const float* grad, long long n, float dt, float dx, float maxdv, float* scratch,
                  float* out, hipStream_t s) {
  const int nb = nblocks(n, 512);
  hipLaunchKernelGGL(k_max_alpha, dim3(nb), dim3(kB), 0, s, vel, grad, n, dt, dx, maxdv, scratch);
  hipLaunchKernelGGL(k_min_partials, dim3(1), dim3(kB), 0, s, scratch, nb, out);
}
void op_update_vel(const float* v, const float* dir, const float* steps, int E, long long n, float tmin, float tmax,
                   float* nv, long long nvstride, hipStream_t s) {
  for (int e0 = 0; e0 < E; e0 += 16) {
    const int k = E - e0 < 16 ? E - e0 : 16;
    Steps st{};
    for (int i = 0; i < k; i++) st.v[i] = steps[e0 + i];
    hipLaunchKernelGGL(k_update_vel, dim3(nblocks(n * k)), dim3(kB), 0, s, v, dir, st, k, n, tmin, tmax,
                       nv + (long long)e0 * nvstride, nvstride);
  }
}
void op_refill(float* a, long long stride, int E, const GridGeom& g, hipStream_t s) {
  hipLaunchKernelGGL(k_refill, dim3(nblocks((long long)g.nx * g.ld * E)), dim3(kB), 0, s, a, stride, E, g);
}
void op_band_pack(const float* field, float* band, int bsize, const GridGeom& g, hipStream_t s) {
  if (bsize > 0) hipLaunchKernelGGL(k_band_pack, dim3((bsize + kB - 1) / kB), dim3(kB), 0, s, field, band, bsize, g);
}
void op_band_unpack(const float* band, float* field, int bsize, const GridGeom& g, hipStream_t s) {
  if (bsize > 0) hipLaunchKernelGGL(k_band_unpack, dim3((bsize + kB - 1) / kB), dim3(kB), 0, s, band, field, bsize, g);
}
void op_reciprocal(const float* vel, float* rv, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_reciprocal2, dim3(nblocks(n)), dim3(kB), 0, s, vel, rv, n);
}
void op_axpy1(const float* src, float* dst, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy1, dim3(nblocks(n)), dim3(kB), 0, s, src, dst, n);
}

namespace {
// One float4 per thread over a grid that covers the array: on MI355X this
// form copies at ~6.2 TB/s, against ~4.6 TB/s for a grid-stride loop over
// 8192 blocks (more independent 1 KiB wavefront loads in flight).
__global__ void k_copy4(const float4* __restrict__ src, float4* __restrict__ dst, long long n4) {
  const long long i = blockIdx.x * (long long)kB + threadIdx.x;
  if (i < n4) dst[i] = src[i];
}
}  // namespace

double op_copy_bandwidth(long long n, int iters, hipStream_t s) {
  const long long n4 = n / 4;
  DBuf<float> a((size_t)n4 * 4), b((size_t)n4 * 4);
  hipLaunchKernelGGL(k_fill_any, dim3(1024), dim3(kB), 0, s, a.get(), n4 * 4);
  hipEvent_t e0, e1;
  SWFWI_HIP(hipEventCreate(&e0));
  SWFWI_HIP(hipEventCreate(&e1));
  if (n4 / kB + 1 > 0x7fffffffLL) fail(kErrArg, "copy size too large");
  const int grid = (int)((n4 + kB - 1) / kB);
  hipLaunchKernelGGL(k_copy4, dim3(grid), dim3(kB), 0, s, (const float4*)a.get(), (float4*)b.get(), n4);
  SWFWI_HIP(hipEventRecord(e0, s));
  for (int i = 0; i < iters; i++)
    hipLaunchKernelGGL(k_copy4, dim3(grid), dim3(kB), 0, s, (const float4*)a.get(), (float4*)b.get(), n4);
  SWFWI_HIP(hipEventRecord(e1, s));
  SWFWI_HIP(hipEventSynchronize(e1));
  float ms = 0;
  SWFWI_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 2.0 * 16.0 * (double)n4 * iters / (ms * 1e-3) / 1e9;
}

namespace {
__global__ void k_touch(float* p) {
  if (threadIdx.x == 0) p[blockIdx.x] = 1.0f;
}
}  // namespace

double op_launch_latency(int blocks, int reps, bool graph, hipStream_t s) {
  DBuf<float> buf((size_t)blocks);
  hipEvent_t e0, e1;
  SWFWI_HIP(hipEventCreate(&e0));
  SWFWI_HIP(hipEventCreate(&e1));
  hipGraphExec_t exec = nullptr;
  if (graph) {
    hipGraph_t g = nullptr;
    SWFWI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < reps; i++) hipLaunchKernelGGL(k_touch, dim3(blocks), dim3(kB), 0, s, buf.get());
    SWFWI_HIP(hipStreamEndCapture(s, &g));
    SWFWI_HIP(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    SWFWI_HIP(hipGraphLaunch(exec, s));  // warm
  } else {
    for (int i = 0; i < 8; i++) hipLaunchKernelGGL(k_touch, dim3(blocks), dim3(kB), 0, s, buf.get());
  }
  SWFWI_HIP(hipEventRecord(e0, s));
  if (graph) {
    SWFWI_HIP(hipGraphLaunch(exec, s));
  } else {
    for (int i = 0; i < reps; i++) hipLaunchKernelGGL(k_touch, dim3(blocks), dim3(kB), 0, s, buf.get());
  }
  SWFWI_HIP(hipEventRecord(e1, s));
  SWFWI_HIP(hipEventSynchronize(e1));
  float ms = 0;
  SWFWI_HIP(hipEventElapsedTime(&ms, e0, e1));
  if (exec) (void)hipGraphExecDestroy(exec);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  SWFWI_HIP(hipGetLastError());
  return 1e3 * ms / reps;
}

void op_gram_f64(const double* A, const double* B, int N, long long m, double* C, hipStream_t s) {
  if (N >= 1 && N <= kGMaxN) {
    // one block per run of rows, two blocks per CU; partials summed in order
    long long nblk = 512;
    long long rows = (m + nblk - 1) / nblk;
    rows = ((rows + kG2KC - 1) / kG2KC) * kG2KC;
    if (rows < kG2KC) rows = kG2KC;
    nblk = (m + rows - 1) / rows;
    if (nblk < 1) nblk = 1;
    DBuf<double> part((size_t)nblk * N * N);
    const bool sym = A == B;
    if (sym)
      hipLaunchKernelGGL(k_gram_rows<true>, dim3((unsigned)nblk), dim3(kG2Threads), 0, s, A, B, N, m, rows,
                         part.get());
    else
      hipLaunchKernelGGL(k_gram_rows<false>, dim3((unsigned)nblk), dim3(kG2Threads), 0, s, A, B, N, m, rows,
                         part.get());
    hipLaunchKernelGGL(k_gram_mfma_finish, dim3(nblocks((long long)N * N)), dim3(kB), 0, s, part.get(), (int)nblk, N,
                       sym, C);
    SWFWI_HIP(hipStreamSynchronize(s));
    return;
  }
  const int nt = (N + kGT - 1) / kGT;
  long long nchunk = 64;
  long long chunk = (m + nchunk - 1) / nchunk;
  chunk = ((chunk + kGK - 1) / kGK) * kGK;
  nchunk = (m + chunk - 1) / chunk;
  if (nchunk < 1) nchunk = 1;
  DBuf<double> part((size_t)nchunk * N * N);
  hipLaunchKernelGGL(k_gram, dim3(nt, nt, (unsigned)nchunk), dim3(kGT * kGT), 0, s, A, B, N, m, chunk, part.get());
  hipLaunchKernelGGL(k_gram_finish, dim3(nblocks((long long)N * N)), dim3(kB), 0, s, part.get(), (int)nchunk, N * N, C);
  SWFWI_HIP(hipStreamSynchronize(s));
}
void op_apply_nxn(const double* Ap, const double* T, int N, long long n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_apply_nxn, dim3(nblocks(n, 2048)), dim3(kB), (size_t)N * N * sizeof(double), s, Ap, T, N, n,
                     out);
}

}  // namespace swfwi
