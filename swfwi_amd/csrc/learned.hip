// Kernels of the learned coarse-grid correction (learned.h).
//
// One block per 16 x 16 tile of output points (z fastest), 256 threads (4
// waves).  Forward, per tile, every layer on BF16 MFMA with the 16 points of
// a group on the product's N side:
//   stage   p and r p (BF16 pairs) with a 3-cell halo (22 x 22) in LDS;
//   layer 1 D1[c][pt] = sum_k W1^T[c][k] X[k][pt], k = 2 t + (p | r p),
//           over the tile + 2 (20 x 20): two 16x16x16 products (K = 18
//           padded to 32), ReLU, BF16 into LDS;
//   layer 2 D2[co][pt] = sum_t sum_ci W2_t^T[co][ci] h1[ci](pt + o_t) over
//           the tile + 1 (18 x 18): five 16x16x32 products (taps in pairs),
//           ReLU, BF16, kept in registers;
//   layer 3 U[t][pt] = sum_c W3[t][c] h2[c](pt) on the same 18 x 18 points,
//           one 16x16x16 product whose B operand is layer 2's result as it
//           stands (its D layout is the B layout), U into LDS over h1's
//           memory once every wave is past layer 2;
//   output  y(x) = sum_t U[t](x + o_t) on the tile.
// So a 16-point group costs 2 + 5 + 1 products and h2 never leaves the
// registers.
//
// v_mfma_f32_16x16x16_bf16 lane layout (measured on the device): lane l
// supplies A[i = l%16][k = 4*(l/16) + v] and B[k = 4*(l/16) + v][j = l%16]
// as its 4 BF16 items v, and holds D[i = 4*(l/16) + v][j = l%16] in
// accumulator item v.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "learned.h"

namespace swfwi {
namespace {

constexpr int kT = 256;
constexpr int kTile = 16;
constexpr int kP = kTile + 6;   // staged field (halo 3)
constexpr int kH1 = kTile + 4;  // layer-1 region (halo 2)
constexpr int kH2 = kTile + 2;  // layer-2 region (halo 1)
constexpr int kNH1 = kH1 * kH1;  // 400
constexpr int kNH2 = kH2 * kH2;  // 324
constexpr int kGroups1 = kNH1 / 16;         // 25 groups of 16 layer-1 points
constexpr int kGroups2 = (kNH2 + 15) / 16;  // 21 groups of 16 layer-2 points
static_assert(kNH1 % 16 == 0, "layer-1 points fill whole groups");

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 hbf16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 hbf16x8 __attribute__((ext_vector_type(8)));

__device__ __host__ __forceinline__ uint16_t bf16_bits(unsigned u) {  // round to nearest even
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ uint16_t bf16_of(float x) { return bf16_bits(__float_as_uint(x)); }
__device__ __forceinline__ float f32_of(uint16_t h) { return __uint_as_float((unsigned)h << 16); }

// accumulator item v of lane l holds D[row][col]
__device__ __forceinline__ int d_row(int l, int v) { return 4 * (l >> 4) + v; }
__device__ __forceinline__ int d_col(int l) { return l & 15; }

// (a, b) rounded to BF16 (nearest even) by the hardware conversion
// (v_cvt_pk_bf16_f32), a in the low half; the same bits as bf16_of for
// finite values
__device__ __forceinline__ unsigned pack_bf16(float a, float b) {
  return __builtin_bit_cast(unsigned, __builtin_convertvector(f32x2{a, b}, hbf16x2));
}

// ReLU then BF16 of four accumulator items.  Rounded first, then the
// negative ones (sign bit set: negative as 16-bit integers, -0 included)
// raised to +0 by one packed 16-bit integer max per pair: the same bits as
// rounding max(x, 0).
__device__ __forceinline__ bf16x4 relu_bf16(const f32x4& a) {
  typedef short s16x2 __attribute__((ext_vector_type(2)));
  const s16x2 zero = {0, 0};
  union { s16x2 h[2]; bf16x4 b; } r;
  r.h[0] = __builtin_elementwise_max(__builtin_bit_cast(s16x2, pack_bf16(a[0], a[1])), zero);
  r.h[1] = __builtin_elementwise_max(__builtin_bit_cast(s16x2, pack_bf16(a[2], a[3])), zero);
  return r.b;
}

// Element i of a field whose base is block-uniform, the byte offset formed in
// 32 bits (a field holds fewer than 2^29 points, checked), so the load takes
// the base from scalar registers and one 32-bit offset
__device__ __forceinline__ float ld_at(const float* base, unsigned i) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + i * 4u);
}

struct TileMap {
  int e, ox, oz;  // entry, tile origin (global padded coords)
};

// tiles over the region grown by `margin` cells on every side
__device__ __forceinline__ TileMap tile_of(const NetGeom& g, int margin) {
  const int tz = (g.z1 - g.z0 + 2 * margin + kTile - 1) / kTile, tx = (g.x1 - g.x0 + 2 * margin + kTile - 1) / kTile;
  const int per = tz * tx;
  TileMap m;
  m.e = blockIdx.x / per;
  const int t = blockIdx.x - m.e * per;
  m.ox = g.x0 - margin + (t / tz) * kTile;
  m.oz = g.z0 - margin + (t % tz) * kTile;
  return m;
}

// Shared forward work of a tile.  sx[kP*kP]: staged (p, r p) as BF16 pairs
// (p in the low half); h1[kNH1*16]: layer 1 (BF16).  Layer 2 then goes to
// h2[kNH2*16] (BF16) when h2 is given and through layer 3 into
// u[9][kNH2] (FP32) when u is given.  Returns false, with nothing computed,
// when `skip_zero` is set and the staged input is zero everywhere (the
// network, ReLU without biases, then gives exactly zero).
__device__ bool forward_tile(const float* __restrict__ p, const float* __restrict__ r, const NetGeom& g,
                             const TileMap& m, const NetDev& w, unsigned* sx, uint16_t* h1, uint16_t* h2, float* u,
                             bool skip_zero) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int nonzero = 0;
  for (int q = tid; q < kP * kP; q += kT) {
    const int x = m.ox - 3 + q / kP, z = m.oz - 3 + q % kP;
    const bool in = x >= 0 && x < g.nx && z >= 0 && z < g.nz;
    const unsigned o = (unsigned)(x * g.nz + z);
    const float v = in ? ld_at(p, o) : 0.f;
    const float vq = in ? ld_at(r, o) * v : 0.f;
    sx[q] = pack_bf16(v, vq);
    nonzero |= v != 0.f;
  }
  const int kq = lane >> 4;  // this lane's K quarter
  const bf16x4 a1a = *reinterpret_cast<const bf16x4*>(w.w1a + lane * 4);
  const bf16x4 a1b = *reinterpret_cast<const bf16x4*>(w.w1a + (64 + lane) * 4);
  if (!__syncthreads_or(nonzero) && skip_zero) return false;  // block-uniform
  // layer 1: lane's B items are (p, r p) at taps 2 kq and 2 kq + 1 (first
  // product) and at tap 8 (second product: k = 16 + 4 kq + v, of which only
  // k = 16, 17 have non-zero weights; the other lanes' finite values add
  // exact zeros)
  {
    const int t0 = 2 * kq, t1 = 2 * kq + 1;
    const int o0 = (t0 / 3 - 1) * kP + (t0 % 3 - 1), o1 = (t1 / 3 - 1) * kP + (t1 % 3 - 1);
    const int o8 = kP + 1;
    constexpr int kMaxG1 = (kGroups1 + kT / 64 - 1) / (kT / 64);  // groups per wavefront, at most
#pragma unroll
    for (int i = 0; i < kMaxG1; i++) {
      const int grp = wave + i * (kT / 64);
      if (grp >= kGroups1) break;  // wave-uniform
      const int q = grp * 16 + (lane & 15);
      const int s = (q / kH1 + 1) * kP + (q % kH1 + 1);  // staged centre
      union { unsigned u2[2]; bf16x4 b; } b0, b1;
      b0.u2[0] = sx[s + o0];
      b0.u2[1] = sx[s + o1];
      b1.u2[0] = sx[s + o8];
      b1.u2[1] = 0u;
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1a, b0.b, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1b, b1.b, acc, 0, 0, 0);
      *reinterpret_cast<bf16x4*>(h1 + q * kNetC + 4 * kq) = relu_bf16(acc);  // channels 4 kq .. 4 kq + 3
    }
  }
  __syncthreads();
  // layers 2 and 3
  // five K = 32 products (v_mfma_f32_16x16x32_bf16: lane l supplies
  // A[i = l%16][k = 8*(l/16) + v], B[k = 8*(l/16) + v][j = l%16]): product q
  // takes tap 2q at k < 16 and tap 2q+1 at k >= 16 (channel k%16); the fifth
  // takes tap 8 at k < 16 and zero weights.  One opcode for the whole accumulation
  // chain: a 16x16x16 product that read the accumulator of a 16x16x32 one
  // issued just before it has given wrong sums on gfx950 (no wait states are
  // inserted between the two).
  hbf16x8 a2p[5];
#pragma unroll
  for (int q = 0; q < 5; q++) a2p[q] = *reinterpret_cast<const hbf16x8*>(w.w2p + (q * 64 + lane) * 8);
  const bf16x4 a3 = *reinterpret_cast<const bf16x4*>(w.w3a + lane * 4);
  const int khalf = kq >> 1, kc = 8 * (kq & 1);  // this lane's tap of the pair, first channel
  // B operand of product q < 4: channels kc .. kc + 7 of tap 2q + khalf of
  // the point's layer-1 neighbourhood; of product 4: those of tap 8 for every
  // lane (the weights past k = 15 are zero and layer-1 values finite, so the
  // lanes of k >= 16 add exact zeros)
  int boff[5];
#pragma unroll
  for (int q = 0; q < 5; q++) {
    const int t = q < 4 ? 2 * q + khalf : 8;
    boff[q] = ((t / 3 - 1) * kH1 + (t % 3 - 1)) * kNetC + kc;
  }
  // layer 2 per group, its result kept in registers; then (apply) layer 3
  // into u, which may share its memory with h1: every layer-2 read of h1 is
  // done at the barrier between the two
  constexpr int kMaxG = (kGroups2 + kT / 64 - 1) / (kT / 64);  // groups per wavefront, at most
  bf16x4 hbs[kMaxG];
#pragma unroll
  for (int i = 0; i < kMaxG; i++) {
    const int grp = wave + i * (kT / 64);
    if (grp >= kGroups2) break;                // wave-uniform
    const int qa = grp * 16 + (lane & 15);     // this lane's point
    const int qs = qa < kNH2 ? qa : kNH2 - 1;  // points past the region read a valid one
    const int x = qs / kH2, z = qs % kH2;      // layer-2 coords; layer-1 = +1
    const uint16_t* c1 = h1 + ((x + 1) * kH1 + (z + 1)) * kNetC;  // the point in layer 1
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 5; q++) {
      const hbf16x8 b = *reinterpret_cast<const hbf16x8*>(c1 + boff[q]);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2p[q], b, acc, 0, 0, 0);
    }
    hbs[i] = relu_bf16(acc);  // h2[4 kq + v][qa]
    if (h2 && qa < kNH2) *reinterpret_cast<bf16x4*>(h2 + qa * kNetC + 4 * kq) = hbs[i];
  }
  if (u) {  // uniform over the grid
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kMaxG; i++) {
      const int grp = wave + i * (kT / 64);
      if (grp >= kGroups2) break;
      const int qa = grp * 16 + (lane & 15);
      f32x4 y3 = {0.f, 0.f, 0.f, 0.f};
      y3 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3, hbs[i], y3, 0, 0, 0);  // U[4 kq + v][qa]
      if (qa < kNH2) {
        float* ub = u + 4 * kq * kNH2 + qa;  // row 4 kq; rows past 8 are not kept
#pragma unroll
        for (int v = 0; v < 4; v++)
          if (4 * kq + v < 9) ub[v * kNH2] = y3[v];
      }
    }
  }
  __syncthreads();
  return true;
}

// 8 waves per SIMD (63 VGPRs, no spills): 5-6% faster than the 7 the
// compiler picks by itself (58 VGPRs plus the accumulators)
__global__ __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(8))) void k_net_apply(const float* __restrict__ p_all, float* __restrict__ y_all,
                                                  NetGeom g, NetDev w) {
  __shared__ unsigned sx[kP * kP];
  // h1 (layer 1) and then u (layer 3): u overwrites h1 after layer 2
  __shared__ __align__(16) uint16_t h1[kNH1 * kNetC];
  static_assert(9 * kNH2 * sizeof(float) <= kNH1 * kNetC * sizeof(uint16_t), "u fits in h1");
  float* u = reinterpret_cast<float*>(h1);
  const TileMap m = tile_of(g, 0);
  const float* p = p_all + (long long)m.e * g.stride;
  const float* r = g.r + (long long)(m.e / g.rdiv) * g.rstride;
  const int x = threadIdx.x / kTile, z = threadIdx.x % kTile;
  const int gx = m.ox + x, gz = m.oz + z;
  float* yo = y_all + (long long)m.e * g.stride + (long long)gx * g.nz + gz;
  const bool out = gx < g.x1 && gz < g.z1;
  if (!forward_tile(p, r, g, m, w, sx, h1, nullptr, u, true)) {
    if (out) *yo = 0.f;
    return;
  }
  if (!out) return;
  float y = 0.f;
#pragma unroll
  for (int t = 0; t < 9; t++) y += u[t * kNH2 + (x + 1 + t / 3 - 1) * kH2 + (z + 1 + t % 3 - 1)];
  *yo = y;
}

__global__ void k_net_add(float* p_all, const float* y_all, NetGeom g) {
  const long long per = (long long)(g.x1 - g.x0) * (g.z1 - g.z0);
  const long long n = per * g.entries;
  for (long long i = blockIdx.x * (long long)kT + threadIdx.x; i < n; i += (long long)gridDim.x * kT) {
    const long long e = i / per, r = i - e * per;
    const int x = g.x0 + (int)(r / (g.z1 - g.z0)), z = g.z0 + (int)(r % (g.z1 - g.z0));
    const long long o = e * g.stride + (long long)x * g.nz + z;
    p_all[o] += y_all[o];
  }
}

__device__ __forceinline__ double block_sum_d(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kT / 64; i++) s += red[i];
  return s;
}

// r = p + y - t over the region; dy = 2 r / count; per-block sum of r^2
__global__ void k_net_loss(const float* p, const float* y, const float* t, float* dy, NetGeom g, double inv_count,
                           double* parts) {
  __shared__ double red[kT / 64];
  const long long per = (long long)(g.x1 - g.x0) * (g.z1 - g.z0);
  const long long n = per * g.entries;
  double acc = 0;
  for (long long i = blockIdx.x * (long long)kT + threadIdx.x; i < n; i += (long long)gridDim.x * kT) {
    const long long e = i / per, r = i - e * per;
    const int x = g.x0 + (int)(r / (g.z1 - g.z0)), z = g.z0 + (int)(r % (g.z1 - g.z0));
    const long long o = e * g.stride + (long long)x * g.nz + z;
    const float res = p[o] + y[o] - t[o];
    dy[o] = (float)(2.0 * res * inv_count);
    acc += (double)res * res;
  }
  const double s = block_sum_d(acc, red);
  if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

// d loss / d weights of one tile: recompute the forward pieces, then
//   dW3[t][c] = sum_tile dy(x) h2[c](x + o_t)
//   dz2[c](x') = relu'(z2) sum_t W3[t][c] dy(x' - o_t)        (tile + 1)
//   dW2[t][ci][co] = sum_tile h1[ci](x + o_t) dz2[co](x)        (MFMA, K = tile points)
//   dz1[ci](x) = relu'(z1) sum_t sum_co W2[t][ci][co] dz2[co](x - o_t)   (MFMA)
//   dW1[t][k][c] = sum_tile in_k(x + o_t) dz1[c](x), in_0 = p, in_1 = r p
// into parts[block][kNetParams].  The forward values enter as the forward
// products saw them: inputs, h1, h2 and W3 rounded to BF16.
__global__ __launch_bounds__(kT) void k_net_backward(const float* __restrict__ p_all, const float* __restrict__ dy_all,
                                                     NetGeom g, NetDev w, float* __restrict__ parts) {
  __shared__ unsigned sx[kP * kP];
  __shared__ __align__(16) uint16_t h1[kNH1 * kNetC];
  __shared__ uint16_t h2[kNH2 * kNetC];
  __shared__ float sdy[kNH1];                 // dy over the tile + 2
  __shared__ uint16_t dz2[kNH2 * kNetC];      // BF16 (MFMA operand)
  __shared__ float dz1[kTile * kTile * kNetC];
  // tiles cover the region grown by 2: dz2 is non-zero up to 1 and dz1 up
  // to 2 cells outside the region, and every point's terms belong to the
  // one tile that owns it
  const TileMap m = tile_of(g, 2);
  const long long eo = (long long)m.e * g.stride;
  forward_tile(p_all + eo, g.r + (long long)(m.e / g.rdiv) * g.rstride, g, m, w, sx, h1, h2, nullptr, false);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int q = tid; q < kNH1; q += kT) {
    const int x = m.ox - 2 + q / kH1, z = m.oz - 2 + q % kH1;
    const bool in = x >= g.x0 && x < g.x1 && z >= g.z0 && z < g.z1;  // dy is zero outside the region
    sdy[q] = in ? dy_all[eo + (long long)x * g.nz + z] : 0.f;
  }
  __syncthreads();
  float* out = parts + (long long)blockIdx.x * kNetParams;
  // dz2 over the tile + 1
  for (int q = tid; q < kNH2; q += kT) {
    const int x = q / kH2, z = q % kH2;  // layer-2 coords; sdy coords = +1
#pragma unroll
    for (int c = 0; c < kNetC; c++) {
      float a = 0.f;
#pragma unroll
      for (int t = 0; t < 9; t++) {
        const int dx = t / 3 - 1, dz = t % 3 - 1;
        a = fmaf(w.w3[t * kNetC + c], sdy[(x + 1 - dx) * kH1 + (z + 1 - dz)], a);
      }
      dz2[q * kNetC + c] = bf16_of(f32_of(h2[q * kNetC + c]) > 0.f ? a : 0.f);
    }
  }
  // dW3: 144 sums over the tile's 256 points, one per thread
  if (tid < kNetW3) {
    const int t = tid / kNetC, c = tid % kNetC;
    const int dx = t / 3 - 1, dz = t % 3 - 1;
    float a = 0.f;
    for (int x = 0; x < kTile; x++)
      for (int z = 0; z < kTile; z++)
        a = fmaf(sdy[(x + 2) * kH1 + (z + 2)], f32_of(h2[((x + 1 + dx) * kH2 + (z + 1 + dz)) * kNetC + c]), a);
    out[kNetW1 + kNetW2 + tid] = a;
  }
  __syncthreads();
  // dW2 on MFMA: per tap, D[ci][co] = sum over the 256 tile points (16 K-steps)
  for (int t = wave; t < 9; t += kT / 64) {
    const int dx = t / 3 - 1, dz = t % 3 - 1;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kc = 0; kc < kTile * kTile / 16; kc++) {
      bf16x4 a, b;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int pt = kc * 16 + 4 * (lane >> 4) + v;  // tile point of this K item
        const int x = pt / kTile, z = pt % kTile;
        a[v] = (short)h1[((x + 2 + dx) * kH1 + (z + 2 + dz)) * kNetC + (lane & 15)];
        b[v] = (short)dz2[((x + 1) * kH2 + (z + 1)) * kNetC + (lane & 15)];
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) out[kNetW1 + (t * kNetC + d_row(lane, v)) * kNetC + d_col(lane)] = acc[v];
  }
  // dz1 on MFMA: per group of 16 tile points, D[pt][ci] = sum_t A_t[pt][co] B_t[co][ci]
  {
    bf16x4 b[9];
#pragma unroll
    for (int t = 0; t < 9; t++) b[t] = *reinterpret_cast<const bf16x4*>(w.w2b + (t * 64 + lane) * 4);
    for (int grp = wave; grp < kTile * kTile / 16; grp += kT / 64) {
      const int pt = grp * 16 + (lane & 15);
      const int x = pt / kTile, z = pt % kTile;
      const int kofs = 4 * (lane >> 4);
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int t = 0; t < 9; t++) {
        const int dx = t / 3 - 1, dz = t % 3 - 1;
        const int q2 = (x + 1 - dx) * kH2 + (z + 1 - dz);
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(dz2 + q2 * kNetC + kofs);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[t], acc, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int pr = grp * 16 + d_row(lane, v), c = d_col(lane);
        const int xr = pr / kTile, zr = pr % kTile;
        const bool on = f32_of(h1[((xr + 2) * kH1 + (zr + 2)) * kNetC + c]) > 0.f;
        dz1[pr * kNetC + c] = on ? acc[v] : 0.f;
      }
    }
  }
  __syncthreads();
  // dW1: 288 sums over the tile
  for (int j = tid; j < kNetW1; j += kT) {
    const int t = j / (kNetIn * kNetC), k = (j / kNetC) % kNetIn, c = j % kNetC;
    const int dx = t / 3 - 1, dz = t % 3 - 1;
    const int sh = 16 * k;  // p in the low half, r p in the high half
    float a = 0.f;
    for (int x = 0; x < kTile; x++)
      for (int z = 0; z < kTile; z++) {
        const float in = f32_of((uint16_t)(sx[(x + 3 + dx) * kP + (z + 3 + dz)] >> sh));
        a = fmaf(in, dz1[(x * kTile + z) * kNetC + c], a);
      }
    out[j] = a;
  }
}

// grad[j] = sum over blocks of parts[b][j] in FP64, in a fixed order: each
// block takes 16 weights; its 16 thread groups sum consecutive runs of the
// partials in block order, then the 16 run sums are added in run order
constexpr int kRedJ = 16, kRedRuns = kT / kRedJ;
__global__ void k_net_reduce(const float* parts, long long nblocks, float* grad) {
  __shared__ double run[kRedRuns][kRedJ];
  const int jl = threadIdx.x % kRedJ, r = threadIdx.x / kRedJ;
  const int j = blockIdx.x * kRedJ + jl;
  const long long per = (nblocks + kRedRuns - 1) / kRedRuns;
  const long long b0 = r * per, b1 = b0 + per < nblocks ? b0 + per : nblocks;
  double s = 0;
  if (j < kNetParams)
    for (long long b = b0; b < b1; b++) s += parts[b * kNetParams + j];
  run[r][jl] = s;
  __syncthreads();
  if (r == 0 && j < kNetParams) {
    double t = 0;
    for (int q = 0; q < kRedRuns; q++) t += run[q][jl];
    grad[j] = (float)t;
  }
}

__global__ void k_net_pack(const float* theta, uint16_t* w1a, uint16_t* w2b, uint16_t* w3a,
                           uint16_t* w2p, float* w3) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i < 5 * 64 * 8) {
    // layer 2, product q: A[co = l%16][k = 8*(l/16) + v] = W2[2q + k/16][k%16][co],
    // zero past tap 8
    const int q = i / 512, l = (i / 8) % 64, v = i % 8;
    const int k = 8 * (l >> 4) + v, co = l & 15, t = 2 * q + (k >> 4);
    w2p[i] = t < 9 ? bf16_of(theta[kNetW1 + (t * kNetC + (k & 15)) * kNetC + co]) : (uint16_t)0;
  }
  if (i < kNetW3) w3[i] = f32_of(bf16_of(theta[kNetW1 + kNetW2 + i]));
  if (i < 2 * 64 * 4) {
    // layer 1, product h: A[c = l%16][k = 16 h + 4*(l/16) + v] = W1 row k (= 2 t + input) or 0 past k = 17
    const int h = i / 256, l = (i / 4) % 64, v = i % 4;
    const int k = 16 * h + 4 * (l >> 4) + v, c = l & 15;
    w1a[i] = k < 9 * kNetIn ? bf16_of(theta[k * kNetC + c]) : (uint16_t)0;
  }
  if (i < 64 * 4) {
    // layer 3: A[t = l%16][c = 4*(l/16) + v] = W3[t][c], rows past tap 8 zero
    const int l = i / 4, v = i % 4;
    const int t = l & 15, c = 4 * (l >> 4) + v;
    w3a[i] = t < 9 ? bf16_of(theta[kNetW1 + kNetW2 + t * kNetC + c]) : (uint16_t)0;
  }
  if (i < 9 * 64 * 4) {
    // item v of lane l for tap t: k = 4*(l/16) + v, n = l%16
    const int t = i / 256, l = (i / 4) % 64, v = i % 4;
    const int k = 4 * (l >> 4) + v, n = l & 15;
    const float* W2 = theta + kNetW1;
    w2b[i] = bf16_of(W2[(t * kNetC + n) * kNetC + k]);  // k = co, n = ci
  }
}

__global__ void k_net_adam(float* theta, const float* grad, float* m, float* v, float c1, float c2, float lr,
                           float b1, float b2, float eps) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= kNetParams) return;
  const float gr = grad[i];
  const float mi = b1 * m[i] + (1.f - b1) * gr;
  const float vi = b2 * v[i] + (1.f - b2) * gr * gr;
  m[i] = mi;
  v[i] = vi;
  theta[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// *out = sum of parts[0..n) in FP64 in a fixed order: consecutive runs per
// thread, then the run sums pairwise in a fixed tree
__global__ void k_sum_parts(const double* parts, int n, double* out) {
  __shared__ double red[kT];
  const int per = (n + kT - 1) / kT, i0 = threadIdx.x * per;
  double s = 0;
  for (int i = i0; i < i0 + per && i < n; i++) s += parts[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = kT / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// the minibatch: sample idx[b] of the training set into slot b, for the
// inputs, models and targets at once
__global__ void k_gather3(float* din, float* dr, float* dt, const float* sin, const float* sr, const float* st,
                          const int* idx, int batch, long long stride) {
  const long long n = (long long)batch * stride;
  for (long long i = blockIdx.x * (long long)kT + threadIdx.x; i < n; i += (long long)gridDim.x * kT) {
    const long long b = i / stride, o = i - b * stride;
    const long long src = (long long)idx[b] * stride + o;
    din[i] = sin[src];
    dr[i] = sr[src];
    dt[i] = st[src];
  }
}

long long tiles_of(const NetGeom& g, int margin) {
  const long long tz = (g.z1 - g.z0 + 2 * margin + kTile - 1) / kTile,
                  tx = (g.x1 - g.x0 + 2 * margin + kTile - 1) / kTile;
  return tz * tx * g.entries;
}

// the staged field reaches 3 cells beyond the tiles, which reach `margin`
// beyond the region
void check_geom(const NetGeom& g, int margin) {
  if (!g.r || g.rdiv < 1) fail(kErrArg, "the network needs the model (1 / vel_t) of every field");
  const int r = 3 + margin;
  if (g.x0 < r || g.z0 < r || g.x1 > g.nx - r || g.z1 > g.nz - r || g.x1 <= g.x0 || g.z1 <= g.z0 || g.entries < 1)
    fail(kErrArg, "network region must lie 5 cells inside the field");
  if (tiles_of(g, margin) > 0x7fffffffLL) fail(kErrRange, "network launch too large");
  if ((long long)g.nx * g.nz >= (1LL << 29)) fail(kErrRange, "network field too large");
}

}  // namespace

void net_pack(const float* theta, const NetDev& w, hipStream_t s) {
  hipLaunchKernelGGL(k_net_pack, dim3((5 * 64 * 8 + kT - 1) / kT), dim3(kT), 0, s, theta, const_cast<uint16_t*>(w.w1a),
                     const_cast<uint16_t*>(w.w2b), const_cast<uint16_t*>(w.w3a),
                     const_cast<uint16_t*>(w.w2p), const_cast<float*>(w.w3));
}

void net_apply(const float* p, float* y, const NetGeom& g, const NetDev& w, hipStream_t s) {
  check_geom(g, 2);
  hipLaunchKernelGGL(k_net_apply, dim3((unsigned)tiles_of(g, 0)), dim3(kT), 0, s, p, y, g, w);
}

void net_add(float* p, const float* y, const NetGeom& g, hipStream_t s) {
  hipLaunchKernelGGL(k_net_add, dim3(2048), dim3(kT), 0, s, p, y, g);
}

void net_loss(const float* p, const float* y, const float* t, float* dy, const NetGeom& g, double* loss_parts,
              int nparts, hipStream_t s) {
  const double count = (double)(g.x1 - g.x0) * (g.z1 - g.z0) * g.entries;
  hipLaunchKernelGGL(k_net_loss, dim3(nparts), dim3(kT), 0, s, p, y, t, dy, g, 1.0 / count, loss_parts);
}

long long net_backward_blocks(const NetGeom& g) { return tiles_of(g, 2); }

void net_backward(const float* p, const float* dy, const NetGeom& g, const NetDev& w, float* parts, float* grad,
                  hipStream_t s) {
  check_geom(g, 2);
  const long long nb = tiles_of(g, 2);
  hipLaunchKernelGGL(k_net_backward, dim3((unsigned)nb), dim3(kT), 0, s, p, dy, g, w, parts);
  hipLaunchKernelGGL(k_net_reduce, dim3((kNetParams + kRedJ - 1) / kRedJ), dim3(kT), 0, s, parts, nb, grad);
}

void net_adam(float* theta, const float* grad, float* m, float* v, int step, float lr, float b1, float b2, float eps,
              hipStream_t s) {
  const float c1 = 1.f / (1.f - std::pow(b1, (float)step)), c2 = 1.f / (1.f - std::pow(b2, (float)step));
  hipLaunchKernelGGL(k_net_adam, dim3((kNetParams + kT - 1) / kT), dim3(kT), 0, s, theta, grad, m, v, c1, c2, lr, b1,
                     b2, eps);
}

void net_reduce_parts(const double* parts, int nparts, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kT), 0, s, parts, nparts, out);
}

// ---- host handle -------------------------------------------------------------

ConvNet::ConvNet() {
  theta_.resize(kNetParams);
  m_.resize(kNetParams);
  v_.resize(kNetParams);
  grad_.resize(kNetParams);
  w3_.resize(kNetW3);
  w1a_.resize(2 * 64 * 4);
  w2b_.resize(9 * 64 * 4);
  w3a_.resize(64 * 4);
  w2p_.resize(5 * 64 * 8);
  lparts_.resize(1024);
  lsum_.resize(1);
}

void ConvNet::repack(hipStream_t s) { net_pack(theta_.get(), dev(), s); }

void ConvNet::set_weights(const float* host) {
  SWFWI_HIP(hipMemcpy(theta_.get(), host, sizeof(float) * kNetParams, hipMemcpyHostToDevice));
  SWFWI_HIP(hipMemset(m_.get(), 0, sizeof(float) * kNetParams));
  SWFWI_HIP(hipMemset(v_.get(), 0, sizeof(float) * kNetParams));
  step_ = 0;
  repack(nullptr);
  SWFWI_HIP(hipDeviceSynchronize());
}

void ConvNet::get_weights(float* host) const {
  SWFWI_HIP(hipMemcpy(host, theta_.get(), sizeof(float) * kNetParams, hipMemcpyDeviceToHost));
}

void ConvNet::correct(float* p, const NetGeom& g, float* yscratch, hipStream_t s) const {
  net_apply(p, yscratch, g, dev(), s);
  net_add(p, yscratch, g, s);
}

// loss into *loss_dev (sum of r^2, FP64) and the weight gradient into grad_
void ConvNet::loss_grad_async(const float* in, const float* target, const NetGeom& g, double* loss_dev,
                              hipStream_t s) {
  const size_t n = (size_t)g.stride * g.entries;
  y_.resize(n);
  dy_.resize(n);
  parts_.resize((size_t)net_backward_blocks(g) * kNetParams);
  net_apply(in, y_.get(), g, dev(), s);
  const int np = 1024;
  net_loss(in, y_.get(), target, dy_.get(), g, lparts_.get(), np, s);
  net_reduce_parts(lparts_.get(), np, loss_dev, s);
  net_backward(in, dy_.get(), g, dev(), parts_.get(), grad_.get(), s);
}

double ConvNet::loss_grad(const float* in, const float* target, const NetGeom& g, float* grad_host, hipStream_t s) {
  loss_grad_async(in, target, g, lsum_.get(), s);
  double loss = 0;
  SWFWI_HIP(hipMemcpyAsync(&loss, lsum_.get(), sizeof(double), hipMemcpyDeviceToHost, s));
  if (grad_host) SWFWI_HIP(hipMemcpyAsync(grad_host, grad_.get(), sizeof(float) * kNetParams, hipMemcpyDeviceToHost, s));
  SWFWI_HIP(hipStreamSynchronize(s));
  const double count = (double)(g.x1 - g.x0) * (g.z1 - g.z0) * g.entries;
  return loss / count;
}

void ConvNet::train_step(const float* in, const float* target, const NetGeom& g, float lr, double* loss_dev,
                         hipStream_t s) {
  loss_grad_async(in, target, g, loss_dev, s);
  step_++;
  net_adam(theta_.get(), grad_.get(), m_.get(), v_.get(), step_, lr, 0.9f, 0.999f, 1e-8f, s);
  repack(s);
}

void ConvNet::set_data(const float* in_host, const float* r_host, const float* target_host, int n,
                       const NetGeom& g) {
  if (n < 1) fail(kErrArg, "empty training set");
  dg_ = g;
  dg_.entries = 1;
  dg_.stride = (long long)g.nx * g.nz;
  dn_ = n;
  const size_t len = (size_t)n * dg_.stride;
  din_.resize(len);
  dr_.resize(len);
  dtgt_.resize(len);
  SWFWI_HIP(hipMemcpy(din_.get(), in_host, sizeof(float) * len, hipMemcpyHostToDevice));
  SWFWI_HIP(hipMemcpy(dr_.get(), r_host, sizeof(float) * len, hipMemcpyHostToDevice));
  SWFWI_HIP(hipMemcpy(dtgt_.get(), target_host, sizeof(float) * len, hipMemcpyHostToDevice));
}

void ConvNet::train(int batch, int steps, float lr, unsigned long long seed, double* losses, hipStream_t s) {
  if (!dn_) fail(kErrState, "no training set");
  if (batch < 1 || batch > dn_) fail(kErrArg, "minibatch size out of range");
  if (steps < 0) fail(kErrArg, "negative step count");
  NetGeom g = dg_;
  g.entries = batch;
  bin_.resize((size_t)batch * g.stride);
  br_.resize((size_t)batch * g.stride);
  btgt_.resize((size_t)batch * g.stride);
  g.r = br_.get();
  g.rstride = g.stride;
  g.rdiv = 1;
  // seeded shuffle of the sample order (splitmix64 + Fisher-Yates), refreshed
  // each pass; the minibatch of step k is idx[k * batch .. (k + 1) * batch)
  std::vector<int> order(dn_), idx((size_t)steps * batch);
  for (int i = 0; i < dn_; i++) order[i] = i;
  unsigned long long st = seed;
  auto next = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    unsigned long long z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  int pos = dn_;
  for (size_t q = 0; q < idx.size(); q++) {
    if (pos == dn_) {
      for (int i = dn_ - 1; i > 0; i--) std::swap(order[i], order[(int)(next() % (unsigned long long)(i + 1))]);
      pos = 0;
    }
    idx[q] = order[pos++];
  }
  if (steps == 0) return;
  DBuf<int> didx;
  DBuf<double> dloss;
  didx.upload(idx.data(), idx.size(), s);
  dloss.resize((size_t)steps);
  const long long total = (long long)batch * g.stride;
  const unsigned gblocks = (unsigned)std::min<long long>((total + kT - 1) / kT, 2048);
  for (int k = 0; k < steps; k++) {
    hipLaunchKernelGGL(k_gather3, dim3(gblocks), dim3(kT), 0, s, bin_.get(), br_.get(), btgt_.get(), din_.get(),
                       dr_.get(), dtgt_.get(), didx.get() + (size_t)k * batch, batch, g.stride);
    train_step(bin_.get(), btgt_.get(), g, lr, dloss.get() + k, s);
  }
  std::vector<double> l((size_t)steps);
  dloss.download(l.data(), l.size(), s);
  SWFWI_HIP(hipStreamSynchronize(s));
  const double count = (double)(g.x1 - g.x0) * (g.z1 - g.z0) * g.entries;
  if (losses)
    for (int k = 0; k < steps; k++) losses[k] = l[k] / count;
}

}  // namespace swfwi
