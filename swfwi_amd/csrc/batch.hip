// Step kernels of the mixed-size batch (batch.h): one launch advances every
// active entry by one time step, each block one 64 (z) x 16 or 32 (x) tile of one
// entry.  Per block: stage the entry's current field with a 6-cell halo in
// LDS (whole rows per wavefront), form the 10th-order Laplacian u2 in LDS
// (each lane sliding its z column over x, the x-direction terms in
// registers), then the damped leapfrog update of fd4t10s-damp-zjh.c:14-115
// in FP32 (fused multiply-adds), each thread 4 or 8 consecutive x at one z with
// its model and previous field loaded before the staging.
//
// Forward (FWD): the next step's source amplitude is added to `next`
// (addSource before stepForward, damp4t10d.cpp:272-289), receivers are
// recorded from the staged current field (recordSeis after the swap), the
// current field is stored as a snapshot on steps 4k+3 with the max-abs scale
// that the previous launch accumulated, and on steps before a snapshot the
// block's max |next| is folded into the next snapshot's scale.
//
// Backward: the receiver field of calgradient (fwiframework.cpp:155-176):
// the next iteration's adjoint source is added to `next` at the receivers,
// and on iterations 4k+4 the imaging condition g -= 4 w(it) s r is formed
// with the source field C[it-1] read back from snapshot k.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "batch.h"
#include "kernels.h"

namespace swfwi {
namespace {

constexpr int kBT = 256;
constexpr int kBZ = kBatchTileZ, kBH = 6;
static_assert(kBZ == 64, "a wavefront covers one z column of the tile");
constexpr int kSZ = kBZ + 2 * kBH;  // staged z
constexpr int kUZ = kBZ + 2;        // u2 z (outputs +-1)
constexpr float kF16Scale = 16384.f;  // snapshot max maps to 2^14 in FP16

// Zhang Jinhai coefficients as the reference stores them (fd4t10s-damp-zjh.c:24-29)
constexpr float kA0 = 1.53400796f, kA1 = 1.78858721f, kA2 = -0.31660756f, kA3 = 0.07612173f, kA4 = -0.01626042f,
                kA5 = 0.00216736f;

__device__ __forceinline__ float damping(int ix, int iz, int nx, int nz, int nb) {
  // fd4t10s-damp-zjh.c:64-80: left, right, then bottom (wins in the corners)
  float dist = 0.f;
  if (ix < nb) dist = (float)(nb - ix) / nb;
  if (ix >= nx - nb) dist = (float)(ix - (nx - nb) + 1) / nb;
  if (iz >= nz - nb) dist = (float)(iz - (nz - nb) + 1) / nb;
  return 0.05f * dist * dist;
}

__device__ __forceinline__ unsigned short to_bf16(float x) {  // round to nearest even
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float from_bf16(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// Element i of an array whose base is block-uniform (scalar registers), the
// byte offset formed in 32 bits: an entry's field, model, checkpoint or
// gradient holds fewer than 2^29 points (checked at creation), so per-point
// accesses need no 64-bit address arithmetic.
template <class T>
__device__ __forceinline__ T& at(T* base, unsigned i) {
  using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return *reinterpret_cast<T*>(reinterpret_cast<B*>(base) + i * (unsigned)sizeof(T));
}

template <int PREC>
__device__ __forceinline__ void snap_store(void* base, unsigned i, float v, float scale) {
  if constexpr (PREC == kSnapF32) {
    at(static_cast<float*>(base), i) = v;
  } else if constexpr (PREC == kSnapBF16) {
    const float inv = scale > 0.f ? 1.f / scale : 0.f;
    at(static_cast<unsigned short*>(base), i) = to_bf16(v * inv);
  } else {
    const float inv = scale > 0.f ? kF16Scale / scale : 0.f;
    at(static_cast<_Float16*>(base), i) = (_Float16)(v * inv);
  }
}
template <int PREC>
__device__ __forceinline__ float snap_load(const void* base, unsigned i, float scale) {
  if constexpr (PREC == kSnapF32) {
    return at(static_cast<const float*>(base), i);
  } else if constexpr (PREC == kSnapBF16) {
    return from_bf16(at(static_cast<const unsigned short*>(base), i)) * scale;
  } else {
    return (float)at(static_cast<const _Float16*>(base), i) * (scale * (1.f / kF16Scale));
  }
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// u2 of the staged field at index c (the unscaled Laplacian of the step)
template <int ORDER>
__device__ __forceinline__ float lap_at(const float* sc, int c) {
  float acc;
  if constexpr (ORDER == kOrder10) {
    acc = -4.f * kA0 * sc[c];
    acc = fmaf(kA1, (sc[c - 1] + sc[c + 1]) + (sc[c - kSZ] + sc[c + kSZ]), acc);
    acc = fmaf(kA2, (sc[c - 2] + sc[c + 2]) + (sc[c - 2 * kSZ] + sc[c + 2 * kSZ]), acc);
    acc = fmaf(kA3, (sc[c - 3] + sc[c + 3]) + (sc[c - 3 * kSZ] + sc[c + 3 * kSZ]), acc);
    acc = fmaf(kA4, (sc[c - 4] + sc[c + 4]) + (sc[c - 4 * kSZ] + sc[c + 4 * kSZ]), acc);
    acc = fmaf(kA5, (sc[c - 5] + sc[c + 5]) + (sc[c - 5 * kSZ] + sc[c + 5 * kSZ]), acc);
  } else {  // 4th-order Laplacian: -5 p + 4/3 (ring 1) - 1/12 (ring 2)
    acc = -5.f * sc[c];
    acc = fmaf(4.f / 3.f, (sc[c - 1] + sc[c + 1]) + (sc[c - kSZ] + sc[c + kSZ]), acc);
    acc = fmaf(-1.f / 12.f, (sc[c - 2] + sc[c + 2]) + (sc[c - 2 * kSZ] + sc[c + 2 * kSZ]), acc);
  }
  return acc;
}

template <int PREC, bool FWD, int ORDER, int TX>
__global__ __launch_bounds__(kBT) void k_batch_step(BatchKernelArgs a) {
  constexpr int kBX = TX;                 // tile x
  constexpr int kPX = kBX / (kBT / kBZ);  // outputs per thread, consecutive in x
  constexpr int kSX = kBX + 2 * kBH;      // staged x
  constexpr int kUX = kBX + 2;            // u2 x (outputs +-1)
  static_assert(kBX % (kBT / kBZ) == 0, "the tile's x splits over the 4 thread rows");
  __shared__ float sc[kSX * kSZ];
  __shared__ float su[kUX * kUZ];
  __shared__ float sm[kBT / 64];
  // XCD-contiguous runs of the work list: blocks are dealt round-robin to the
  // 8 XCDs, so consecutive tiles of one entry share an XCD's L2
  const int n = a.nblocks, per = n / 8, rem = n % 8;
  const int xcd = blockIdx.x % 8, local = blockIdx.x / 8;
  const int l = xcd * per + (xcd < rem ? xcd : rem) + local;
  const int pk = a.blk[l];
  const int e = pk >> kBlkTileBits;
  const BatchEntry E = a.ent[e];
  const int tile = pk & ((1 << kBlkTileBits) - 1);
  const int it = FWD ? a.it_or_k : E.nt - 1 - a.it_or_k;
  const int tx = tile / E.ntz, tz = tile - tx * E.ntz;
  const int ox = kBH + tx * kBX, oz = kBH + tz * kBZ;
  const int ld = E.nzpad;
  const long long npts = (long long)E.nxpad * ld;
  const bool even = (a.rot & 1) == 0;
  const float* cur = (even ? a.fb : a.fa) + E.f;
  float* nxt = (even ? a.fa : a.fb) + E.f;  // holds prev; next is written over it

  const bool inside = ox - kBH + kSX <= E.nxpad && oz - kBH + kSZ <= E.nzpad;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // this thread's kPX outputs (consecutive x at one z): the model and prev
  // are loaded before the staging so that their latency overlaps it
  const int zl = threadIdx.x & (kBZ - 1), xg = threadIdx.x / kBZ;
  const int gz = oz + zl;
  const float* rvb = a.rv + E.model;
  float rvj[kPX], pvj[kPX];
#pragma unroll
  for (int j = 0; j < kPX; j++) {
    const int gx = ox + xg * kPX + j;
    const bool ok = gx < E.nxpad - kBH && gz < E.nzpad - kBH;
    const int p = gx * ld + gz;
    rvj[j] = ok ? at(rvb, p) : 0.f;
    pvj[j] = ok ? at(nxt, p) : 0.f;
  }
  // backward imaging steps: the checkpoint and the gradient of the thread's
  // outputs are loaded here too, ahead of the staging
  const bool image_now =
      !FWD && PREC != kSnapNone && it >= kSnapEvery && (it % kSnapEvery == 0) && a.wts[E.wts0 + it] != 0.f;
  const long long snap_off = E.snap + (long long)(FWD ? it / kSnapEvery : it / kSnapEvery - 1) * npts;
  void* snapb = PREC == kSnapNone ? a.snap
                : PREC == kSnapF32 ? static_cast<void*>(static_cast<float*>(a.snap) + snap_off)
                                   : static_cast<void*>(static_cast<unsigned short*>(a.snap) + snap_off);
  float* gradb = a.grad + E.grad;
  float svj[kPX], gvj[kPX];
  if (image_now) {
    const float iscale = __uint_as_float(a.scale[E.scale0 + it / kSnapEvery - 1]);
#pragma unroll
    for (int j = 0; j < kPX; j++) {
      const int gx = ox + xg * kPX + j;
      const bool ok = gx < E.nxpad - kBH && gz < E.nzpad - kBH;
      const int p = gx * ld + gz;
      svj[j] = ok ? snap_load<PREC == kSnapNone ? kSnapF32 : PREC>(snapb, p, iscale) : 0.f;
      gvj[j] = ok ? at(gradb, p) : 0.f;
    }
  }
  // stage cur + 6-cell halo: each wavefront takes whole rows (one z per
  // lane, the last 12 z of a row by lanes 0-11), so a row is two coalesced
  // loads and no lane divides its index
  for (int xx = wave; xx < kSX; xx += kBT / 64) {
    const int gx = ox - kBH + xx;
    const unsigned src = gx * ld + (oz - kBH) + lane;
    float v0 = 0.f, v1 = 0.f;
    if (gx < E.nxpad) {  // wave-uniform
      if (inside || oz - kBH + lane < E.nzpad) v0 = at(cur, src);
      if (lane < kSZ - 64 && (inside || oz - kBH + 64 + lane < E.nzpad)) v1 = at(cur, src + 64);
    }
    sc[xx * kSZ + lane] = v0;
    if (lane < kSZ - 64) sc[xx * kSZ + 64 + lane] = v1;
  }
  __syncthreads();
  // u2 on the tile + 1 ring: each wavefront takes a run of rows (kRows or
  // kRows - 1) and holds its lane's staged column (tile z = lane) over those
  // rows and their x reach in registers, so the x-direction terms are read
  // from LDS once per row instead of once per tap; the loop is unrolled over
  // the rows, so the window never shifts.  The ring's z = -1 and z = 64
  // points follow, one per thread of waves 0-1.  The terms are summed in the
  // same order as lap_at below.
  {
    constexpr int R = ORDER == kOrder10 ? 5 : 2;
    constexpr int kW = kBT / 64;                // wavefronts
    constexpr int kRows = (kUX + kW - 1) / kW;  // most u2 rows of a wavefront
    const int r0 = wave * kUX / kW, nrows = (wave + 1) * kUX / kW - r0;  // wave-uniform
    // staged rows r0 - 1 - R + k of the column of tile z = lane, k < kRows +
    // 2R (a wavefront with kRows - 1 rows reads one row it does not use: it
    // is not the last wavefront, which takes kRows, so that row lies inside
    // the staged tile); every read is an immediate offset from one register
    int base = lane + kBH + (r0 + kBH - 1 - R) * kSZ;
    asm("" : "+v"(base));
    const float* col = sc + base;
    float w[kRows + 2 * R];
#pragma unroll
    for (int k = 0; k < kRows + 2 * R; k++) w[k] = col[k * kSZ];
#pragma unroll
    for (int i = 0; i < kRows; i++) {
      if (i < kRows - 1 || nrows == kRows) {
        const int xu = r0 + i;
        // staged row of u2 row xu, its z reads as immediate offsets 0..2R
        // from one register (opaque, so the reads are not re-based in pairs)
        int ro = (i + R) * kSZ - R;
        asm("" : "+v"(ro));
        const float* row = col + ro + R;
        const float* x = w + i;  // x[R]: the row's own staged value
        float acc;
        if constexpr (ORDER == kOrder10) {
          acc = -4.f * kA0 * x[5];
          acc = fmaf(kA1, (row[-1] + row[1]) + (x[4] + x[6]), acc);
          acc = fmaf(kA2, (row[-2] + row[2]) + (x[3] + x[7]), acc);
          acc = fmaf(kA3, (row[-3] + row[3]) + (x[2] + x[8]), acc);
          acc = fmaf(kA4, (row[-4] + row[4]) + (x[1] + x[9]), acc);
          acc = fmaf(kA5, (row[-5] + row[5]) + (x[0] + x[10]), acc);
        } else {
          acc = -5.f * x[2];
          acc = fmaf(4.f / 3.f, (row[-1] + row[1]) + (x[1] + x[3]), acc);
          acc = fmaf(-1.f / 12.f, (row[-2] + row[2]) + (x[0] + x[4]), acc);
        }
        su[xu * kUZ + lane + 1] = acc;
      }
    }
    if (threadIdx.x < 2 * kUX) {
      const int xu = threadIdx.x >> 1, zu = (threadIdx.x & 1) ? kUZ - 1 : 0;
      su[xu * kUZ + zu] = lap_at<ORDER>(sc, (xu + kBH - 1) * kSZ + zu + kBH - 1);
    }
  }
  __syncthreads();

  const int rz = E.rec0 % ld, rx0 = E.rec0 / ld;
  // the sponge (fd4t10s-damp-zjh.c:64-80) reaches this tile only near the
  // left, right or bottom edge; elsewhere d = 0 and (2-2d+d^2, 1-2d) = (2, 1)
  const bool sponge = ox < E.damp || ox + kBX > E.nxpad - E.damp || oz + kBZ > E.nzpad - E.damp;
  const bool snap_now = FWD && PREC != kSnapNone && (it % kSnapEvery == kSnapEvery - 1);
  const bool max_now = FWD && ((it + 1) % kSnapEvery == kSnapEvery - 1) && (it + 1) / kSnapEvery < E.nsnap;
  const bool inject_next = FWD ? it + 1 < E.nt : it - 1 >= E.it_last;
  float scale = 0.f;
  if (snap_now) scale = __uint_as_float(a.scale[E.scale0 + it / kSnapEvery]);
  if (image_now) scale = __uint_as_float(a.scale[E.scale0 + it / kSnapEvery - 1]);
  const float wimg = image_now ? a.wts[E.wts0 + it] : 0.f;
  float vmax = 0.f;
#pragma unroll
  for (int j = 0; j < kPX; j++) {
    const int xl = xg * kPX + j;
    const int gx = ox + xl;
    if (gx >= E.nxpad - kBH || gz >= E.nzpad - kBH) continue;
    const int p = gx * ld + gz;
    const int uc = (xl + 1) * kUZ + zl + 1;
    const float u = su[uc];
    const float r = rvj[j];
    const float c = sc[(xl + kBH) * kSZ + zl + kBH];
    const float pv = pvj[j];
    float v;
    if (sponge) {
      const float d = E.dscale * damping(gx, gz, E.nxpad, E.nzpad, E.damp);
      v = fmaf(2.f - 2.f * d + d * d, c, -(1.f - 2.f * d) * pv);
    } else {
      v = fmaf(2.f, c, -pv);  // d = 0: the coefficients are exactly 2 and 1
    }
    v = fmaf(r, u, v);
    if constexpr (ORDER == kOrder10) {  // the 4th-order-in-time term r^2/12 Lap5(u2)
      const float lap = (su[uc - 1] + su[uc + 1]) + (su[uc - kUZ] + su[uc + kUZ]) - 4.f * u;
      v = fmaf(r * r * (1.f / 12.f), lap, v);
    }
    const bool rec = gz == rz && gx >= rx0 && gx < rx0 + E.ng;
    if constexpr (FWD) {
      if (inject_next && p == E.src) v += a.amp[E.amp0 + it + 1];
      vmax = fmaxf(vmax, fabsf(v));
      if (rec && a.dcal) a.dcal[E.dcal + (long long)it * E.ng + (gx - rx0)] = c;
      if constexpr (PREC != kSnapNone) {
        if (snap_now) snap_store<PREC>(snapb, p, c, scale);
      }
      if (a.tap_k >= 0 && ((gx | (gz - E.tap_oz)) & 1) == 0) {  // a coarse point
        const int xc = (gx + kBH) >> 1, zc = (gz + kBH - E.tap_oz) >> 1;
        if (xc < E.tap_nx && zc < E.tap_nz)
          a.tap[E.tap + (long long)a.tap_k * E.tap_nx * E.tap_nz + xc * E.tap_nz + zc] = c;
      }
    } else {
      if (inject_next && rec) v += a.dcal[E.dcal + (long long)(it - 1) * E.ng + (gx - rx0)];
      if (image_now && PREC != kSnapNone) at(gradb, p) = gvj[j] - wimg * svj[j] * c;
    }
    at(nxt, p) = v;
  }
  if (max_now) {  // the next step's field is snapshot (it+1)/4: fold its max
    vmax = wave_max(vmax);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = vmax;
    __syncthreads();
    if (threadIdx.x == 0) {
      float m = sm[0];
      for (int w = 1; w < kBT / 64; w++) m = fmaxf(m, sm[w]);
      // non-negative floats order like their bit patterns
      atomicMax(&a.scale[E.scale0 + (it + 1) / kSnapEvery], __float_as_uint(m));
    }
  }
}

// step-0 fields: both buffers zero (by the caller), then the first source
// amplitude (forward) or the last adjoint-source sample (backward) in fb
__global__ void k_batch_inject0(BatchKernelArgs a, int nentries, int fwd) {
  const int e = blockIdx.x;
  if (e >= nentries) return;
  const BatchEntry E = a.ent[e];
  if (fwd) {
    if (threadIdx.x == 0) a.fb[E.f + E.src] = a.amp[E.amp0];
  } else {
    const int ld = E.nzpad;
    for (int r = threadIdx.x; r < E.ng; r += blockDim.x)
      a.fb[E.f + E.rec0 + (long long)r * ld] = a.dcal[E.dcal + (long long)(E.nt - 1) * E.ng + r];
  }
}

__global__ void k_sub(const float* a, const float* b, float* out, long long n) {
  for (long long i = blockIdx.x * (long long)kBT + threadIdx.x; i < n; i += (long long)gridDim.x * kBT)
    out[i] = a[i] - b[i];
}

__global__ void k_fill_u32(unsigned* p, long long n, unsigned v) {
  for (long long i = blockIdx.x * (long long)kBT + threadIdx.x; i < n; i += (long long)gridDim.x * kBT) p[i] = v;
}

}  // namespace

void launch_fill_u32(unsigned* p, long long n, unsigned v, hipStream_t s) {
  long long b = (n + kBT - 1) / kBT;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)b), dim3(kBT), 0, s, p, n, v);
}

void launch_sub(const float* a, const float* b, float* out, long long n, hipStream_t s) {
  long long blk = (n + kBT - 1) / kBT;
  if (blk > 8192) blk = 8192;
  if (blk < 1) blk = 1;
  hipLaunchKernelGGL(k_sub, dim3((unsigned)blk), dim3(kBT), 0, s, a, b, out, n);
}

void launch_batch_init_fwd(const BatchKernelArgs& a, int nentries, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_inject0, dim3(nentries), dim3(64), 0, s, a, nentries, 1);
}
void launch_batch_init_bwd(const BatchKernelArgs& a, int nentries, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_inject0, dim3(nentries), dim3(kBT), 0, s, a, nentries, 0);
}

template <bool FWD, int ORDER, int TX>
static void launch_step(const BatchKernelArgs& a, int prec, hipStream_t s) {
  if (a.nblocks <= 0) return;
  switch (prec) {
    case kSnapF32:
      hipLaunchKernelGGL((k_batch_step<kSnapF32, FWD, ORDER, TX>), dim3(a.nblocks), dim3(kBT), 0, s, a);
      break;
    case kSnapBF16:
      hipLaunchKernelGGL((k_batch_step<kSnapBF16, FWD, ORDER, TX>), dim3(a.nblocks), dim3(kBT), 0, s, a);
      break;
    case kSnapF16:
      hipLaunchKernelGGL((k_batch_step<kSnapF16, FWD, ORDER, TX>), dim3(a.nblocks), dim3(kBT), 0, s, a);
      break;
    default:
      hipLaunchKernelGGL((k_batch_step<kSnapNone, FWD, ORDER, TX>), dim3(a.nblocks), dim3(kBT), 0, s, a);
      break;
  }
}
template <bool FWD, int ORDER>
static void launch_width(const BatchKernelArgs& a, int prec, int tile_x, hipStream_t s) {
  if (tile_x == 32)
    launch_step<FWD, ORDER, 32>(a, prec, s);
  else if (tile_x == 24)
    launch_step<FWD, ORDER, 24>(a, prec, s);
  else
    launch_step<FWD, ORDER, 16>(a, prec, s);
}
template <bool FWD>
static void launch_dir(const BatchKernelArgs& a, int prec, int order, int tile_x, hipStream_t s) {
  if (order == kOrder4)
    launch_width<FWD, kOrder4>(a, prec, tile_x, s);
  else
    launch_width<FWD, kOrder10>(a, prec, tile_x, s);
}
void launch_batch_fwd(const BatchKernelArgs& a, int prec, int order, int tile_x, hipStream_t s) {
  launch_dir<true>(a, prec, order, tile_x, s);
}
void launch_batch_bwd(const BatchKernelArgs& a, int prec, int order, int tile_x, hipStream_t s) {
  launch_dir<false>(a, prec, order, tile_x, s);
}

}  // namespace swfwi
