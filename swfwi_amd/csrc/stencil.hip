// The damped/undamped 4t10s leapfrog step as fused, LDS-tiled CDNA4 kernels.
//
// One launch advances every entry of a batch (shots, ensemble members, line
// search trials) by one time step.  Around the stencil it folds in the
// per-step bookkeeping that the reference does in separate host loops:
//   forward  (Damp4t10d::FwiForwardModeling/EssForwardModeling, calgradient's
//            forward half, damp4t10d.cpp:272-307, fwiframework.cpp:137-144):
//            record cur at receivers, save cur's boundary band, inject the next
//            step's source amplitudes into next;
//   backward (calgradient's backward half, fwiframework.cpp:146-176): undamped
//            reconstruction of the source field + damped receiver field in the
//            same tile, subSource, the zero-lag imaging condition, and the
//            band restore / receiver injection that the next iteration needs.
//
// A tile is kTZ x TX output points (TX = 8, 16 or 32, picked per batch
// size by the propagator's measured rule).  Its `cur` plus a 6-cell halo is staged in
// LDS with every global load of the block in flight at once, the unscaled
// 10th-order Laplacian u2 is formed in LDS for the tile plus a 1-cell ring,
// and each thread finishes TX/4 consecutive points along x at its z (the
// wavefront of thread row r the rows r*TX/4 ...).  prev, the model and
// (backward) the gradient of the thread's points are prefetched into
// registers before the staging so their latency overlaps it.
//
// Block order: blocks are spread round-robin over the 8 XCDs, so the block
// index is remapped to give each XCD one contiguous run of the work list.
// Inside a run either all entries of a tile come back to back (the tile's
// shared model is re-read from that XCD's L2), or, for batches with one
// model per entry and for the fused backward, chunks of tile columns come
// entry by entry (a tile's x-halo, the neighbouring columns of the same
// entry, is re-read from L2); the launcher picks the order.
//
// Precision: kExact reproduces the reference's expression trees
// (fd4t10s-damp-zjh.c:44-54,85-88; fd4t10s-zjh.c:28-38,49-52): float ring
// sums, float coefficient products, double accumulation of -4.0*a0*p and the
// (2-2d+d^2)p / 2p terms, one rounding to float on store, no contraction.  The
// library is compiled with -ffp-contract=off; kFp32 uses explicit fmaf.
#include "kernels.h"

#include <type_traits>

namespace swfwi {
namespace {

constexpr int kRows = kThreads / kTZ;  // thread rows per block (4)

// Element i of p, with the byte offset formed in 32 bits (a field holds fewer
// than 2^30 points, checked at handle creation): a load or store then takes
// the block-uniform base from scalar registers and one 32-bit offset per
// lane, with no 64-bit address arithmetic per access.
template <class T>
__device__ __forceinline__ T& elem(T* p, unsigned i) {
  using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return *reinterpret_cast<T*>(reinterpret_cast<B*>(p) + i * (unsigned)sizeof(T));
}

template <int TX>
struct Tile {
  static constexpr int LZ = kTZ + 2 * kHalo;  // LDS row length of the staged cur tile
  static constexpr int LX = TX + 2 * kHalo;
  static constexpr int UZ = kTZ + 2;          // LDS row length of the u2 tile
  static constexpr int UX = TX + 2;
  static constexpr int PER = TX / kRows;      // outputs per thread
  // Staging: every staged row's first 64 z (one per lane of the row's
  // wavefront, SROWS rows per thread) and its last 2*kHalo z (SREST per thread).
  static constexpr int ZREST = LZ - kTZ;
  static constexpr int SROWS = LX / kRows;
  static constexpr int SREST = (ZREST * LX + kThreads - 1) / kThreads;
  static constexpr int NSTAGE = SROWS + SREST;
  // u2: the tile's own 64 z of every u2 row (a run of consecutive rows per
  // wavefront, one row more in the first UX % kRows wavefronts), and the two
  // ring z of every row (2*UX points, taken by the wavefronts with one row
  // fewer).
  static_assert(LX % kRows == 0, "staged rows split evenly over the wavefronts");
  static_assert(2 * UX <= kThreads - (UX % kRows) * kTZ, "ring points fit the wavefronts with fewer rows");
};

// Zhang Jinhai coefficients as float, as the reference stores them.
constexpr float A0 = 1.53400796f, A1 = 1.78858721f, A2 = -0.31660756f;
constexpr float A3 = 0.07612173f, A4 = -0.01626042f, A5 = 0.00216736f;
constexpr double M4A0 = -4.0 * (double)A0;
constexpr float C12 = 1.0f / 12;

// u2 from the centre value and the five ring sums r_k = ((z-k + z+k) + x-k)
// + x+k, summed as the reference does.
template <int P>
__device__ __forceinline__ float u2_sum(float s0, float r1, float r2, float r3, float r4, float r5) {
  if constexpr (P == kExact) {
    double acc = M4A0 * (double)s0;
    acc = acc + (double)(A1 * r1);
    acc = acc + (double)(A2 * r2);
    acc = acc + (double)(A3 * r3);
    acc = acc + (double)(A4 * r4);
    acc = acc + (double)(A5 * r5);
    return (float)acc;
  } else {
    float acc = (-4.0f * A0) * s0;
    acc = __builtin_fmaf(A1, r1, acc);
    acc = __builtin_fmaf(A2, r2, acc);
    acc = __builtin_fmaf(A3, r3, acc);
    acc = __builtin_fmaf(A4, r4, acc);
    acc = __builtin_fmaf(A5, r5, acc);
    return acc;
  }
}

template <int P, int LZ>
__device__ __forceinline__ float u2_at(const float* __restrict__ s) {
  // s: LDS element; z neighbours at +-k, x neighbours at +-k*LZ
  const float r1 = ((s[-1] + s[1]) + s[-LZ]) + s[LZ];
  const float r2 = ((s[-2] + s[2]) + s[-2 * LZ]) + s[2 * LZ];
  const float r3 = ((s[-3] + s[3]) + s[-3 * LZ]) + s[3 * LZ];
  const float r4 = ((s[-4] + s[4]) + s[-4 * LZ]) + s[4 * LZ];
  const float r5 = ((s[-5] + s[5]) + s[-5 * LZ]) + s[5 * LZ];
  return u2_sum<P>(s[0], r1, r2, r3, r4, r5);
}

// The same with the x neighbours taken from w[0..10] = the column at x-5..x+5.
template <int P>
__device__ __forceinline__ float u2_window(const float* __restrict__ s, const float (&w)[11]) {
  const float r1 = ((s[-1] + s[1]) + w[4]) + w[6];
  const float r2 = ((s[-2] + s[2]) + w[3]) + w[7];
  const float r3 = ((s[-3] + s[3]) + w[2]) + w[8];
  const float r4 = ((s[-4] + s[4]) + w[1]) + w[9];
  const float r5 = ((s[-5] + s[5]) + w[0]) + w[10];
  return u2_sum<P>(w[5], r1, r2, r3, r4, r5);
}

// next from cur c, prev p, reciprocal model rv, the u2 value of the point
// u0 and its z (zm, zp) and x (xm, xp) neighbours, damping delta (0 inside;
// ignored when !DAMPED).
template <bool DAMPED, int P>
__device__ __forceinline__ float update_vals(float c, float p, float rv, float u0, float zm, float zp, float xm,
                                             float xp, float delta) {
  float lap = ((zm + zp) + xm) + xp;
  lap = lap - 4 * u0;
  const float cc = (C12 * rv) * rv;
  if constexpr (P == kExact) {
    double acc;
    if constexpr (DAMPED) {
      const double a = (2. - (double)(2 * delta)) + (double)(delta * delta);
      const float b = 1 - 2 * delta;
      acc = a * (double)c;
      acc = acc - (double)(b * p);
    } else {
      acc = 2. * (double)c;
      acc = acc - (double)p;
    }
    acc = acc + (double)(rv * u0);
    acc = acc + (double)(cc * lap);
    return (float)acc;
  } else {
    float a = 2.0f, b = 1.0f;
    if constexpr (DAMPED) {
      a = (2.0f - 2 * delta) + delta * delta;
      b = 1 - 2 * delta;
    }
    float acc = __builtin_fmaf(a, c, -(b * p));
    acc = __builtin_fmaf(rv, u0, acc);
    return __builtin_fmaf(cc, lap, acc);
  }
}

// The u2 column of a thread's outputs: the thread's z, the rows of its PER
// consecutive outputs and one more on either side, read from the u2 tile
// once (x neighbours of one output are the centres of the next and previous).
template <int TX>
__device__ __forceinline__ void u2_column(const float* __restrict__ u2, int row, int lz,
                                          float (&col)[Tile<TX>::PER + 2]) {
  using T = Tile<TX>;
#pragma unroll
  for (int j = 0; j < T::PER + 2; j++) col[j] = u2[(row * T::PER + j) * T::UZ + lz + 1];
}
// next of output k of the thread: u points at its u2 value (z neighbours at
// u[-1], u[1]), col is the thread's u2 column.
template <bool DAMPED, int P, int N>
__device__ __forceinline__ float update_col(float c, float p, float rv, const float (&col)[N], int k,
                                            const float* __restrict__ u, float delta) {
  return update_vals<DAMPED, P>(c, p, rv, col[k + 1], u[-1], u[1], col[k], col[k + 2], delta);
}

// Offset of (ix, iz) in the reference's boundary band (writeBndry layout,
// damp4t10d.cpp:478-514), or -1 when the point is not in the band.
__device__ __forceinline__ int band_index(const Grid& g, int ix, int iz) {
  const int w = 6;
  const int xl = g.bx0 - w;                          // first band column
  const int nxb = g.nx - (g.bx0 - w + g.bxn - w);
  const int nzb = g.nz - g.bz0 - g.bzn;
  const int zb = g.nz - g.bzn;                       // first bottom band row
  if (iz >= zb && iz < zb + w && ix >= xl && ix < xl + nxb) return (iz - zb) + w * (ix - xl);
  if (iz >= g.bz0 && iz < g.bz0 + nzb) {
    if (ix >= xl && ix < g.bx0) return w * nxb + (iz - g.bz0) + nzb * (ix - xl);
    const int xr = g.nx - g.bxn;
    if (ix >= xr && ix < xr + w) return w * nxb + (iz - g.bz0) + nzb * (ix - xr + w);
  }
  return -1;
}

// The damping delta of a point, in halves: the z table entry of the
// thread's row (one per thread), the x table entry of a column (uniform over
// a wavefront, whose 64 lanes share ix: a scalar load through the constant
// address space, the tables being read-only while the kernels run), and the
// pick at the use, so nothing waits on the loads while others are issued.
// Indices are clamped into the tables: a point past the grid edge of a
// partial tile reads an entry it never uses.
__device__ __forceinline__ float delta_z_load(const Grid& g, int iz) { return g.delta_z[min(iz, g.nz - 1)]; }
__device__ __forceinline__ float delta_x_load(const Grid& g, int ix) {
  using ConstFloat = __attribute__((address_space(4))) const float;
  return ((ConstFloat*)g.delta_x)[min(ix, g.nx - 1)];
}
__device__ __forceinline__ float delta_pick(const Grid& g, int iz, float dz, float dx) {
  return iz >= g.zdamp ? dz : dx;
}

// Whether any point of [x0, x1) x [z0, z1) lies in the sponge (delta != 0:
// ix < nb, ix >= nx - nb, iz >= nz - nb).  Where delta is 0 the damped update
// reduces exactly to the undamped one: (2 - 2*0) + 0*0 = 2 and 1 - 2*0 = 1.
__device__ __forceinline__ bool touches_sponge(const Grid& g, int x0, int x1, int z0, int z1) {
  return x0 < g.nb || x1 > g.nx - g.nb || z1 > g.zdamp;
}

// Divisions of the block-index arithmetic, as a multiply-high by
// m = ceil(2^32 / d): exact while numerator x d <= 2^32, which the launcher
// checks for the largest numerator (m = 0 otherwise: plain division), so a
// block starts with a few scalar multiplies instead of the generic integer
// division sequences.
struct Div {
  int d = 1;
  unsigned m = 0;
};
__device__ __forceinline__ int div_of(int n, const Div& v) {
  if (v.d == 1) return n;
  return v.m ? (int)__umulhi((unsigned)n, v.m) : n / v.d;
}
// What a block needs to find its tile and entry: the entry count of the work
// list, the chunk length (block order below), and the divisors by entries,
// chunk x entries, chunk, tiles along z and entries per model.
struct BlockMap {
  int nentries = 1, chunk = 0;
  Div ent, span, chk, ntz, rvdiv;
};

// Block -> (tile, entry): XCD-contiguous runs of the work list.  Not
// CHUNKED: entries fastest (all entries of a tile back to back).  CHUNKED:
// the list walks chunks of `chunk` consecutive tiles (a few tile columns),
// each chunk entry by entry, so that a tile's x-halo (the neighbouring
// columns of the same entry) and, when the entries share a model, the
// chunk's model are re-read from the XCD's L2 instead of HBM.
template <bool CHUNKED>
__device__ __forceinline__ void block_work(int ntiles, const BlockMap& m, int& tile, int& e) {
  const int n = ntiles * m.nentries;
  const int b = blockIdx.x;
  const int per = n / 8, rem = n % 8;
  const int xcd = b % 8, local = b / 8;
  const int l = xcd * per + (xcd < rem ? xcd : rem) + local;
  if constexpr (CHUNKED) {
    const int c = div_of(l, m.span);
    const int r = l - c * m.span.d;
    const int tc = min(m.chunk, ntiles - c * m.chunk);  // the last chunk may be short
    e = tc == m.chunk ? div_of(r, m.chk) : r / tc;
    tile = c * m.chunk + (r - e * tc);
  } else {
    tile = div_of(l, m.ent);
    e = l - tile * m.nentries;
  }
}

// Stage cur (tile + halo) into LDS, zero outside the grid, in two halves so
// that other loads can be issued between them: stage_load issues every load
// of the thread (no branch: a slot past the staged region or outside the grid
// reads element 0 and drops it), stage_store writes them to LDS.  FULL: the
// staged region lies inside the grid (every tile but the last row and
// column), no checks.  The wavefront of thread row r loads the first 64 z of
// staged rows r, r + 4, ... (one z per lane, so every address and LDS offset
// of the pass is a constant step from the first); the last 2*kHalo z of every
// staged row are spread over all threads.
template <int TX>
struct Staged {
  float v[Tile<TX>::NSTAGE];
  int rest[Tile<TX>::SREST];  // LDS index of each of the thread's last-z elements (-1: none)
};
template <int TX, bool FULL>
__device__ __forceinline__ void stage_load(const Grid& g, const float* __restrict__ src, int x0, int z0,
                                           Staged<TX>& st) {
  using T = Tile<TX>;
  const int lane = threadIdx.x % kTZ, row = threadIdx.x / kTZ;
  const int gz = z0 - kHalo + lane;  // >= 0: z0 >= kHalo
#pragma unroll
  for (int k = 0; k < T::SROWS; k++) {
    const int gx = x0 - kHalo + row + kRows * k;  // >= 0: x0 >= kHalo
    const bool ok = FULL || (gx < g.nx && gz < g.nz);
    const float v = elem(src, ok ? (unsigned)(gx * g.ld + gz) : 0u);
    st.v[k] = ok ? v : 0.f;
  }
#pragma unroll
  for (int k = 0; k < T::SREST; k++) {
    const int t = threadIdx.x + k * kThreads;
    const int lx = t / T::ZREST, lz = kTZ + (t - lx * T::ZREST);
    const int gx = x0 - kHalo + lx, gzr = z0 - kHalo + lz;
    const bool in = t < T::ZREST * T::LX;
    const bool ok = in && (FULL || (gx < g.nx && gzr < g.nz));
    const float v = elem(src, ok ? (unsigned)(gx * g.ld + gzr) : 0u);
    st.v[T::SROWS + k] = ok ? v : 0.f;
    st.rest[k] = in ? lx * T::LZ + lz : -1;
  }
}
template <int TX>
__device__ __forceinline__ void stage_store(const Staged<TX>& st, float* __restrict__ lds) {
  using T = Tile<TX>;
  const int lane = threadIdx.x % kTZ, row = threadIdx.x / kTZ;
#pragma unroll
  for (int k = 0; k < T::SROWS; k++) lds[(row + kRows * k) * T::LZ + lane] = st.v[k];
#pragma unroll
  for (int k = 0; k < T::SREST; k++)
    if (st.rest[k] >= 0) lds[st.rest[k]] = st.v[T::SROWS + k];
}

// u2 over the tile plus a one-cell ring, only where the reference defines it
// ([5, n-5) in both directions); elsewhere 0 (never read by a valid output).
// The wavefront of thread row r forms the tile's own 64 z of u2 rows r, r + 4,
// ... (one z per lane); the ring z (one below, one above the tile) of every
// row go to the wavefronts that have one row fewer.
template <int P, int TX, bool FULL>
__device__ __forceinline__ void u2_point(const Grid& g, const float* __restrict__ cur, float* __restrict__ u2,
                                         int x0, int z0, int ux, int uz) {
  using T = Tile<TX>;
  const int gx = x0 - 1 + ux, gz = z0 - 1 + uz;
  float v = 0.f;
  if (FULL || (gx >= 5 && gx < g.nx - 5 && gz >= 5 && gz < g.nz - 5))
    v = u2_at<P, T::LZ>(cur + (ux + kHalo - 1) * T::LZ + (uz + kHalo - 1));
  u2[ux * T::UZ + uz] = v;
}
template <int P, int TX, bool FULL>
__device__ __forceinline__ void form_u2(const Grid& g, const float* __restrict__ cur, float* __restrict__ u2,
                                        int x0, int z0) {
  using T = Tile<TX>;
  const int lane = threadIdx.x % kTZ, row = threadIdx.x / kTZ;
  // the wavefront of thread row r takes the consecutive u2 rows [u0, u0 + n),
  // sliding the column of its lane over x so that each staged value is read
  // from LDS once per wavefront for the x-direction terms
  constexpr int BASE = T::UX / kRows, EXTRA = T::UX % kRows;
  const int n = BASE + (row < EXTRA ? 1 : 0);
  const int u0 = row * BASE + (row < EXTRA ? row : EXTRA);
  const float* __restrict__ col = cur + u0 * T::LZ + lane + kHalo;  // staged row u0 (= x of u2 row u0, minus 5)
  float w[11];
#pragma unroll
  for (int d = 1; d < 11; d++) w[d] = col[(d - 1) * T::LZ];
  const int gz = z0 + lane;
  auto step = [&](int k) {
#pragma unroll
    for (int d = 0; d < 10; d++) w[d] = w[d + 1];
    w[10] = col[(k + 10) * T::LZ];
    const int ux = u0 + k;
    const int gx = x0 - 1 + ux;
    // the row's offset passes through an empty asm so that its z reads are
    // addressed from one register with small immediate offsets (paired LDS
    // reads), not re-based from the thread's first row for every pair
    int row_off = (k + 5) * T::LZ;
    asm("" : "+v"(row_off));
    float v = 0.f;
    if (FULL || (gx >= 5 && gx < g.nx - 5 && gz >= 5 && gz < g.nz - 5)) v = u2_window<P>(col + row_off, w);
    u2[ux * T::UZ + lane + 1] = v;
  };
#pragma unroll
  for (int k = 0; k < BASE; k++) step(k);
  if (n > BASE) step(BASE);  // uniform over the wavefront
  const int t = (int)threadIdx.x - EXTRA * kTZ;
  if (t >= 0 && t < 2 * T::UX) u2_point<P, TX, FULL>(g, cur, u2, x0, z0, t >> 1, (t & 1) ? kTZ + 1 : 0);
}

// Whether the tile may hold points of the boundary band (a superset test;
// band_index decides per point).
template <int TX>
__device__ __forceinline__ bool touches_band(const Grid& g, int x0, int z0) {
  return x0 < g.bx0 || x0 + TX > g.nx - g.bxn || z0 + kTZ > g.nz - g.bzn;
}

// Whether the tile's staged region (tile + 6 cells) lies inside the grid:
// then every load, u2 point and output of the tile is in range.
template <int TX>
__device__ __forceinline__ bool tile_is_full(const Grid& g, int x0, int z0) {
  return x0 + TX + kHalo <= g.nx && z0 + kTZ + kHalo <= g.nz;
}

__device__ __forceinline__ bool entry_active(const Inject& in, int e, int pid) {
  return in.lo == nullptr || (pid >= in.lo[e] && pid < in.hi[e]);
}

// A tile's points of interest staged in LDS at kernel start, so no global
// load depends on the stencil result: grid index (-1 = inactive for this
// entry) and, for injections, the amplitude of this step.  Tiles with more
// than kTilePts points fall back to scanning the global list.
struct TilePts {
  int q[kTilePts];
  int pid[kTilePts];
  float amp[kTilePts];
};

// One thread's share of a tile's point list: threads 0..n-1 hold one item
// each (n is the tile's item count, uniform over the block): grid index (-1
// when inactive for this entry), id, amplitude, position in the tile and
// whether another point of the tile shares the grid index.
struct PtsLoad {
  int n, q, pid, loc, shared;
  float amp;
};
constexpr PtsLoad kNoPts{0, -1, 0, 0, 1, 0.f};

// A slot list is read in two levels so that its latency hides behind the
// field loads.  fetch_slots, before them: the tile's count and one slot per
// thread (threads past kTilePts re-read slot threadIdx % kTilePts and drop
// it; empty slots hold grid index -1 and id 0), and the entry's id range.
// finish_slots, once the field loads are in flight: the staged item and, for
// an injection, its amplitude at this step (only when the tile has points,
// so the amplitude table exists; an empty slot reads point 0's and drops it).
struct SlotFetch {
  int n, lo, hi;
  int4 sl;
};
__device__ __forceinline__ SlotFetch fetch_slots(const PointList& pts, const Inject* in, int tile, int e) {
  SlotFetch f;
  f.n = pts.count[tile];
  f.sl = pts.slots[(long long)tile * kTilePts + (threadIdx.x & (kTilePts - 1))];
  f.lo = 0;
  f.hi = 0x7fffffff;
  if (in && in->lo) {
    f.lo = in->lo[e];
    f.hi = in->hi[e];
  }
  return f;
}
__device__ __forceinline__ PtsLoad finish_slots(const SlotFetch& f, const Inject* in, int e) {
  PtsLoad r;
  r.n = f.n;
  r.pid = f.sl.y;
  r.loc = f.sl.z;
  r.shared = f.sl.w;
  const bool mine = (int)threadIdx.x < f.n;
  r.q = mine && f.sl.y >= f.lo && f.sl.y < f.hi ? f.sl.x : -1;
  r.amp = 0.f;
  if (in && f.n > 0) r.amp = elem(in->amp + e * in->estride + (long long)in->t * in->ld, (unsigned)f.sl.y);
  return r;
}

// A tile's items from the CSR list (tiles of more than kTilePts items make
// the whole list use it): offsets, then ids, then grid indices and
// amplitudes, each level waiting on the one before.
__device__ __forceinline__ PtsLoad load_points(const PointList& pts, const Inject* in, int tile, int e, int x0, int z0,
                                               int ld) {
  PtsLoad r = kNoPts;
  const int b = pts.tile_ptr[tile];
  r.n = pts.tile_ptr[tile + 1] - b;
  if (r.n <= kTilePts && (int)threadIdx.x < r.n) {
    r.pid = pts.items[b + threadIdx.x];
    r.q = pts.idx[r.pid];
    const int ix = r.q / ld;
    r.loc = (ix - x0) * kTZ + (r.q - ix * ld - z0);
  }
  if (in && r.n <= kTilePts && (int)threadIdx.x < r.n) {
    if (entry_active(*in, e, r.pid))
      r.amp = in->amp[e * in->estride + (long long)in->t * in->ld + r.pid];
    else
      r.q = -1;
  }
  return r;
}

// A tile's point list for one launch: slots when the list has them (two
// levels around the field loads), else the CSR scan (at the start).  LEAN:
// the launch guarantees that every active list has slots (no CSR code).
struct TileList {
  bool on, slots;
  SlotFetch f;
  PtsLoad r;
};
template <bool LEAN>
__device__ __forceinline__ TileList list_begin(const PointList& pts, const Inject* in, bool on, int tile, int e,
                                               int x0, int z0, int ld) {
  TileList l;
  l.on = on;
  l.slots = on && (LEAN || pts.slots != nullptr);
  l.r = kNoPts;
  if (l.slots)
    l.f = fetch_slots(pts, in, tile, e);
  else if (!LEAN && on)
    l.r = load_points(pts, in, tile, e, x0, z0, ld);
  return l;
}
__device__ __forceinline__ PtsLoad list_end(const TileList& l, const Inject* in, int e) {
  return l.slots ? finish_slots(l.f, in, e) : l.r;
}

// Grid index and id go to LDS before the first barrier; the amplitude (the
// second level of loads) after it.
__device__ __forceinline__ void store_points(const PtsLoad& r, TilePts& tp) {
  if (r.n <= kTilePts && (int)threadIdx.x < r.n) {
    tp.q[threadIdx.x] = r.q;
    tp.pid[threadIdx.x] = r.pid;
  }
}
__device__ __forceinline__ void store_amp(const PtsLoad& r, TilePts& tp) {
  if (r.n <= kTilePts && (int)threadIdx.x < r.n) tp.amp[threadIdx.x] = r.amp;
}
// Per-tile map of staged injection points: for each output point of the
// tile, the first active slot whose point lands there, with bit 6 set when
// later slots land there too (-1: none).  mark_clear: all threads, before the
// barrier that publishes the staged slots; mark_slots: slot threads, after it.
template <int TX>
__device__ __forceinline__ void mark_clear(signed char* map) {
  for (int i = threadIdx.x; i < kTZ * TX / 4; i += kThreads) reinterpret_cast<int*>(map)[i] = -1;
}
__device__ __forceinline__ void mark_slots(const PtsLoad& r, const TilePts& tp, signed char* map) {
  const int j = threadIdx.x;
  if (j >= r.n || r.q < 0) return;
  if (!r.shared) {  // the only point of the tile at its grid index
    map[r.loc] = (signed char)j;
    return;
  }
  for (int k = 0; k < j; k++)
    if (tp.q[k] == r.q) return;  // an earlier active slot heads this point's chain
  int more = 0;
  for (int k = j + 1; k < r.n; k++)
    if (tp.q[k] == r.q) more = 0x40;
  map[r.loc] = (signed char)(j | more);
}
// Add or subtract, in slot (point id) order, the staged amplitudes landing on
// q, given the map entry m of its output point.
__device__ __forceinline__ float apply_mapped(const TilePts& tp, int n, int m, int sign, int q, float v) {
  const int h = m & 0x3F;
  if (!(m & 0x40)) return sign > 0 ? v + tp.amp[h] : v - tp.amp[h];
  for (int j = h; j < n; j++)
    if (tp.q[j] == q) v = sign > 0 ? v + tp.amp[j] : v - tp.amp[j];
  return v;
}
// Record the staged receivers of the tile from the staged field (slot threads).
__device__ __forceinline__ void record_slots(const PtsLoad& r, const float* tile, int ldz, float* out) {
  if ((int)threadIdx.x >= r.n) return;
  elem(out, (unsigned)r.pid) = tile[((r.loc >> 6) + kHalo) * ldz + (r.loc & (kTZ - 1)) + kHalo];
}

// Apply all active point amplitudes of this tile that land on grid index q,
// in point order (manipSource, damp4t10d.cpp:204-214).
__device__ __forceinline__ float apply_points(const Inject& in, int tile, int e, int q, float v) {
  const int b = in.pts.tile_ptr[tile], eidx = in.pts.tile_ptr[tile + 1];
  for (int j = b; j < eidx; j++) {
    const int pid = in.pts.items[j];
    if (in.pts.idx[pid] == q && entry_active(in, e, pid)) {
      const float a = in.amp[e * in.estride + (long long)in.t * in.ld + pid];
      v = in.sign > 0 ? v + a : v - a;
    }
  }
  return v;
}

template <bool DAMPED, int P, int TX, bool FULL, bool LEAN>
__device__ __forceinline__ void fwd_tile(const FwdStepArgs& a, int tile, int e, int model, int x0, int z0, float* s_cur,
                                         float* s_u2, TilePts& s_inj, TilePts& s_rec, signed char* imap) {
  using T = Tile<TX>;
  const Grid& g = a.g;
  const long long fo = (long long)e * a.fstride;
  float* __restrict__ prev = a.prev + fo;
  const float* __restrict__ rv = a.rv + (long long)model * a.rvstride;

  const int lz = threadIdx.x % kTZ, row = __builtin_amdgcn_readfirstlane(threadIdx.x / kTZ);
  const int iz = z0 + lz;
  // every global load of the block is issued here, before the first barrier:
  // the point lists' first level, the thread's own points (a point past the
  // grid edge of a partial tile reads element 0 and drops it), the staged
  // tile, then the lists' second level.  (The compiler copies the receiver
  // list's loaded slot where the branch that skips an inactive list joins,
  // and that copy waits for the slot load; a branch-free form measured no
  // faster.)
  const TileList ti = list_begin<LEAN>(a.inj.pts, &a.inj, a.inj.t >= 0, tile, e, x0, z0, g.ld);
  const TileList tr = list_begin<LEAN>(a.rec.pts, nullptr, a.rec.out != nullptr, tile, e, x0, z0, g.ld);
  const bool sponge = DAMPED && touches_sponge(g, x0, x0 + TX, z0, z0 + kTZ);
  const bool band_tile = a.band != nullptr && touches_band<TX>(g, x0, z0);
  // (values of points past the grid edge are loaded from element 0 and
  // never used: the update loop skips those points)
  float pv[T::PER], rr[T::PER], dx[T::PER] = {}, dz = 0.f;
  if (sponge) dz = delta_z_load(g, iz);
#pragma unroll
  for (int k = 0; k < T::PER; k++) {
    const int ix = x0 + row * T::PER + k;
    const bool ok = FULL || (ix < g.nx - kHalo && iz < g.nz - kHalo);
    const unsigned q = ok ? (unsigned)(ix * g.ld + iz) : 0u;
    pv[k] = elem(prev, q);
    rr[k] = elem(rv, q);
    if (sponge) dx[k] = delta_x_load(g, ix);
  }
  Staged<TX> st;
  stage_load<TX, FULL>(g, a.cur + fo, x0, z0, st);
  const PtsLoad li = list_end(ti, &a.inj, e), lr = list_end(tr, nullptr, e);
  stage_store<TX>(st, s_cur);
  store_points(li, s_inj);
  store_points(lr, s_rec);
  const bool mapped = li.n > 0 && (LEAN || li.n <= kTilePts);
  if (mapped) mark_clear<TX>(imap);
  __syncthreads();
  if (mapped) {
    store_amp(li, s_inj);
    mark_slots(li, s_inj, imap);
  }
  if (lr.n > 0 && lr.n <= kTilePts)
    record_slots(lr, s_cur, T::LZ, a.rec.out + (long long)e * a.rec.estride);
  form_u2<P, TX, FULL>(g, s_cur, s_u2, x0, z0);
  __syncthreads();

  float ucol[T::PER + 2];
  u2_column<TX>(s_u2, row, lz, ucol);
#pragma unroll
  for (int k = 0; k < T::PER; k++) {
    const int lx = row * T::PER + k;
    const int ix = x0 + lx;
    if (!FULL && (ix >= g.nx - kHalo || iz >= g.nz - kHalo)) continue;
    const unsigned q = (unsigned)(ix * g.ld + iz);
    const float c = s_cur[(lx + kHalo) * T::LZ + lz + kHalo];
    const float* u = s_u2 + (lx + 1) * T::UZ + lz + 1;
    if (!LEAN && a.u2out) a.u2out[fo + q] = u[0];
    float nxt = sponge ? update_col<true, P>(c, pv[k], rr[k], ucol, k, u, delta_pick(g, iz, dz, dx[k]))
                       : update_col<false, P>(c, pv[k], rr[k], ucol, k, u, 0.f);
    if (mapped) {
      const int h = imap[lx * kTZ + lz];
      if (h >= 0) nxt = apply_mapped(s_inj, li.n, h, a.inj.sign, q, nxt);
    } else if (!LEAN && li.n > 0) {
      nxt = apply_points(a.inj, tile, e, q, nxt);
    }
    elem(prev, q) = nxt;
    if (!LEAN && lr.n > kTilePts) {
      const int b = a.rec.pts.tile_ptr[tile], eidx = a.rec.pts.tile_ptr[tile + 1];
      for (int j = b; j < eidx; j++) {
        const int pid = a.rec.pts.items[j];
        if (a.rec.pts.idx[pid] == (int)q) a.rec.out[(long long)e * a.rec.estride + pid] = c;
      }
    }
  }
  // the band of cur (writeBndry), as a pass of its own so that no index
  // arithmetic of it is computed ahead in the tiles that hold no band point
  if (band_tile) {
#pragma unroll
    for (int k = 0; k < T::PER; k++) {
      const int lx = row * T::PER + k;
      const int ix = x0 + lx;
      if (!FULL && (ix >= g.nx - kHalo || iz >= g.nz - kHalo)) continue;
      const int bi = band_index(g, ix, iz);
      if (bi >= 0) elem(a.band + (long long)e * a.band_estride, bi) = s_cur[(lx + kHalo) * T::LZ + lz + kHalo];
    }
  }
  // u2 of the ring [5,6) and [n-6,n-5) that no output point owns: the
  // reference writes it too, so the optional u2 export covers it.
  if (!LEAN && a.u2out) {
    for (int i = threadIdx.x; i < T::UX * T::UZ; i += kThreads) {
      const int ux = i / T::UZ, uz = i - ux * T::UZ;
      const int gx = x0 - 1 + ux, gz = z0 - 1 + uz;
      // ring points may sit in two tiles' u2 halos; both write the same value
      const bool ring = gx == 5 || gz == 5 || gx == g.nx - 6 || gz == g.nz - 6;
      if (ring && gx >= 5 && gx < g.nx - 5 && gz >= 5 && gz < g.nz - 5)
        a.u2out[fo + (long long)gx * g.ld + gz] = s_u2[i];
    }
  }
}

template <bool DAMPED, int P, int TX, bool LEAN, bool CHUNKED>
__global__ __launch_bounds__(kThreads) void k_fwd_step(FwdStepArgs a, BlockMap m) {
  using T = Tile<TX>;
  __shared__ float s_cur[T::LX * T::LZ];
  __shared__ float s_u2[T::UX * T::UZ];
  __shared__ TilePts s_inj, s_rec;
  __shared__ __align__(4) signed char imap[kTZ * TX];
  int tile, e;
  block_work<CHUNKED>(a.g.ntz * a.g.ntx, m, tile, e);
  const int tx = div_of(tile, m.ntz), model = div_of(e, m.rvdiv);
  const int z0 = kHalo + (tile - tx * a.g.ntz) * kTZ, x0 = kHalo + tx * TX;
  if (tile_is_full<TX>(a.g, x0, z0))
    fwd_tile<DAMPED, P, TX, true, LEAN>(a, tile, e, model, x0, z0, s_cur, s_u2, s_inj, s_rec, imap);
  else
    fwd_tile<DAMPED, P, TX, false, LEAN>(a, tile, e, model, x0, z0, s_cur, s_u2, s_inj, s_rec, imap);
}

template <int P, int TX, bool FULL, bool LEAN>
__device__ __forceinline__ void bwd_tile(const BwdStepArgs& a, int tile, int e, int model, int x0, int z0, float* s_sc,
                                         float* s_gc, float* s_su, float* s_gu, TilePts& s_sub, TilePts& s_inj,
                                         signed char* smap, signed char* gmap) {
  using T = Tile<TX>;
  const Grid& g = a.g;
  const long long fo = (long long)e * a.fstride;
  const float* __restrict__ rv = a.rv + (long long)model * a.rvstride;
  float* __restrict__ s_prev = a.s_prev + fo;
  float* __restrict__ g_prev = a.g_prev + fo;
  float* __restrict__ grad = a.grad + (long long)e * a.gstride;

  const int lz = threadIdx.x % kTZ, row = __builtin_amdgcn_readfirstlane(threadIdx.x / kTZ);
  const int iz = z0 + lz;
  // every global load of the block is issued here, before the first barrier,
  // and no branch waits on one (as in the forward tile)
  const TileList ts = list_begin<LEAN>(a.s_sub.pts, &a.s_sub, a.s_sub.t >= 0, tile, e, x0, z0, g.ld);
  const TileList tg = list_begin<LEAN>(a.g_inj.pts, &a.g_inj, a.g_inj.t >= 0, tile, e, x0, z0, g.ld);
  const bool sponge = touches_sponge(g, x0, x0 + TX, z0, z0 + kTZ);
  const bool band_tile = a.band != nullptr && touches_band<TX>(g, x0, z0);
  // (values of points past the grid edge are loaded from element 0 and
  // never used: the update loop skips those points; bv is used only in band)
  float sp[T::PER], gp[T::PER], rr[T::PER], gr[T::PER] = {}, dx[T::PER] = {}, bv[T::PER] = {}, dz = 0.f, pf[T::PER] = {};
  unsigned inband = 0;
  if (sponge) dz = delta_z_load(g, iz);
#pragma unroll
  for (int k = 0; k < T::PER; k++) {
    const int ix = x0 + row * T::PER + k;
    const bool ok = FULL || (ix < g.nx - kHalo && iz < g.nz - kHalo);
    const unsigned q = ok ? (unsigned)(ix * g.ld + iz) : 0u;
    sp[k] = elem(s_prev, q);
    gp[k] = elem(g_prev, q);
    rr[k] = elem(rv, q);
    if (a.do_xcorr) gr[k] = elem(grad, q);
    if (sponge) dx[k] = delta_x_load(g, ix);
    if (band_tile) {
      const int bi = ok ? band_index(g, ix, iz) : -1;
      bv[k] = elem(a.band + (long long)e * a.band_estride, bi >= 0 ? (unsigned)bi : 0u);
      // the next launch's band, read ahead so that it comes from the
      // memory-side cache rather than HBM (the values are not used)
      if (a.band_next) pf[k] = elem(a.band_next + (long long)e * a.band_estride, bi >= 0 ? (unsigned)bi : 0u);
      if (bi >= 0) inband |= 1u << k;
    }
  }
  Staged<TX> st_s, st_g;
  stage_load<TX, FULL>(g, a.s_cur + fo, x0, z0, st_s);
  stage_load<TX, FULL>(g, a.g_cur + fo, x0, z0, st_g);
  const PtsLoad ls = list_end(ts, &a.s_sub, e), li = list_end(tg, &a.g_inj, e);
  stage_store<TX>(st_s, s_sc);
  stage_store<TX>(st_g, s_gc);
  store_points(ls, s_sub);
  store_points(li, s_inj);
  const bool smapped = ls.n > 0 && (LEAN || ls.n <= kTilePts), gmapped = li.n > 0 && (LEAN || li.n <= kTilePts);
  if (smapped) mark_clear<TX>(smap);
  if (gmapped) mark_clear<TX>(gmap);
  __syncthreads();
  if (smapped) {
    store_amp(ls, s_sub);
    mark_slots(ls, s_sub, smap);
  }
  if (gmapped) {
    store_amp(li, s_inj);
    mark_slots(li, s_inj, gmap);
  }
  form_u2<P, TX, FULL>(g, s_sc, s_su, x0, z0);
  form_u2<P, TX, FULL>(g, s_gc, s_gu, x0, z0);
  __syncthreads();
  float scol[T::PER + 2], gcol[T::PER + 2];
  u2_column<TX>(s_su, row, lz, scol);
  u2_column<TX>(s_gu, row, lz, gcol);

#pragma unroll
  for (int k = 0; k < T::PER; k++) {
    const int lx = row * T::PER + k;
    const int ix = x0 + lx;
    if (!FULL && (ix >= g.nx - kHalo || iz >= g.nz - kHalo)) continue;
    const unsigned q = (unsigned)(ix * g.ld + iz);
    const int l = (lx + kHalo) * T::LZ + lz + kHalo;
    const int lu = (lx + 1) * T::UZ + lz + 1;
    // source field: stepBackward (undamped) then subSource
    float sn = update_col<false, P>(s_sc[l], sp[k], rr[k], scol, k, s_su + lu, 0.f);
    if (smapped) {
      const int h = smap[lx * kTZ + lz];
      if (h >= 0) sn = apply_mapped(s_sub, ls.n, h, a.s_sub.sign, q, sn);
    } else if (!LEAN && ls.n > 0) {
      sn = apply_points(a.s_sub, tile, e, q, sn);
    }
    // receiver field: stepForward (damped) of the already injected cur
    const float gc = s_gc[l];
    float gn = sponge ? update_col<true, P>(gc, gp[k], rr[k], gcol, k, s_gu + lu, delta_pick(g, iz, dz, dx[k]))
                      : update_col<false, P>(gc, gp[k], rr[k], gcol, k, s_gu + lu, 0.f);
    // zero-lag imaging condition with the new source field and the injected
    // receiver field (cross_correlation, fwiframework.cpp:107-112)
    if (a.do_xcorr) elem(grad, q) = gr[k] - sn * gc * a.weight;
    // what the next iteration (it-1) expects to find: the band of the source
    // field restored (readBndry) and the receivers injected (addSource)
    if (inband & (1u << k)) sn = bv[k];
    if (gmapped) {
      const int h = gmap[lx * kTZ + lz];
      if (h >= 0) gn = apply_mapped(s_inj, li.n, h, a.g_inj.sign, q, gn);
    } else if (!LEAN && li.n > 0) {
      gn = apply_points(a.g_inj, tile, e, q, gn);
    }
    elem(s_prev, q) = sn;
    elem(g_prev, q) = gn;
  }
  if (band_tile && a.band_next) {  // keep the read-ahead loads (nothing reads their values)
#pragma unroll
    for (int k = 0; k < T::PER; k++) asm volatile("" ::"v"(pf[k]));
  }
}

// The same backward iteration as two independent halves of one launch, for
// small batches whose fused tiles leave most of the chip idle: the source
// half (undamped step, subSource, imaging condition with the receiver field
// read at the point, band restore) and the receiver half (damped step,
// receiver injection).  Each half stages one field.
template <int P, int TX, bool FULL, bool LEAN>
__device__ __forceinline__ void bwd_src_tile(const BwdStepArgs& a, int tile, int e, int model, int x0, int z0, float* s_c,
                                             float* s_u, TilePts& tp, signed char* map) {
  using T = Tile<TX>;
  const Grid& g = a.g;
  const long long fo = (long long)e * a.fstride;
  const float* __restrict__ rv = a.rv + (long long)model * a.rvstride;
  const float* __restrict__ g_cur = a.g_cur + fo;
  float* __restrict__ s_prev = a.s_prev + fo;
  float* __restrict__ grad = a.grad + (long long)e * a.gstride;
  const int lz = threadIdx.x % kTZ, row = __builtin_amdgcn_readfirstlane(threadIdx.x / kTZ);
  const int iz = z0 + lz;
  // loads issued as in the forward tile: no branch waits on one
  const TileList ts = list_begin<LEAN>(a.s_sub.pts, &a.s_sub, a.s_sub.t >= 0, tile, e, x0, z0, g.ld);
  const bool band_tile = a.band != nullptr && touches_band<TX>(g, x0, z0);
  // (values of points past the grid edge are loaded from element 0 and
  // never used: the update loop skips those points; bv is used only in band)
  float sp[T::PER], rr[T::PER], gr[T::PER] = {}, gc[T::PER] = {}, bv[T::PER] = {}, pf[T::PER] = {};
  unsigned inband = 0;
#pragma unroll
  for (int k = 0; k < T::PER; k++) {
    const int ix = x0 + row * T::PER + k;
    const bool ok = FULL || (ix < g.nx - kHalo && iz < g.nz - kHalo);
    const unsigned q = ok ? (unsigned)(ix * g.ld + iz) : 0u;
    sp[k] = elem(s_prev, q);
    rr[k] = elem(rv, q);
    if (a.do_xcorr) {
      gr[k] = elem(grad, q);
      gc[k] = elem(g_cur, q);
    }
    if (band_tile) {
      const int bi = ok ? band_index(g, ix, iz) : -1;
      bv[k] = elem(a.band + (long long)e * a.band_estride, bi >= 0 ? (unsigned)bi : 0u);
      // the next launch's band, read ahead so that it comes from the
      // memory-side cache rather than HBM (the values are not used)
      if (a.band_next) pf[k] = elem(a.band_next + (long long)e * a.band_estride, bi >= 0 ? (unsigned)bi : 0u);
      if (bi >= 0) inband |= 1u << k;
    }
  }
  Staged<TX> st;
  stage_load<TX, FULL>(g, a.s_cur + fo, x0, z0, st);
  const PtsLoad ls = list_end(ts, &a.s_sub, e);
  stage_store<TX>(st, s_c);
  store_points(ls, tp);
  const bool mapped = ls.n > 0 && (LEAN || ls.n <= kTilePts);
  if (mapped) mark_clear<TX>(map);
  __syncthreads();
  if (mapped) {
    store_amp(ls, tp);
    mark_slots(ls, tp, map);
  }
  form_u2<P, TX, FULL>(g, s_c, s_u, x0, z0);
  __syncthreads();
  float ucol[T::PER + 2];
  u2_column<TX>(s_u, row, lz, ucol);
#pragma unroll
  for (int k = 0; k < T::PER; k++) {
    const int lx = row * T::PER + k;
    const int ix = x0 + lx;
    if (!FULL && (ix >= g.nx - kHalo || iz >= g.nz - kHalo)) continue;
    const unsigned q = (unsigned)(ix * g.ld + iz);
    float sn = update_col<false, P>(s_c[(lx + kHalo) * T::LZ + lz + kHalo], sp[k], rr[k], ucol, k,
                                    s_u + (lx + 1) * T::UZ + lz + 1, 0.f);
    if (mapped) {
      const int h = map[lx * kTZ + lz];
      if (h >= 0) sn = apply_mapped(tp, ls.n, h, a.s_sub.sign, q, sn);
    } else if (!LEAN && ls.n > 0) {
      sn = apply_points(a.s_sub, tile, e, q, sn);
    }
    if (a.do_xcorr) elem(grad, q) = gr[k] - sn * gc[k] * a.weight;
    if (inband & (1u << k)) sn = bv[k];
    elem(s_prev, q) = sn;
  }
  if (band_tile && a.band_next) {  // keep the read-ahead loads (nothing reads their values)
#pragma unroll
    for (int k = 0; k < T::PER; k++) asm volatile("" ::"v"(pf[k]));
  }
}

template <int P, int TX, bool FULL, bool LEAN>
__device__ __forceinline__ void bwd_rcv_tile(const BwdStepArgs& a, int tile, int e, int model, int x0, int z0, float* s_c,
                                             float* s_u, TilePts& tp, signed char* map) {
  using T = Tile<TX>;
  const Grid& g = a.g;
  const long long fo = (long long)e * a.fstride;
  const float* __restrict__ rv = a.rv + (long long)model * a.rvstride;
  float* __restrict__ g_prev = a.g_prev + fo;
  const int lz = threadIdx.x % kTZ, row = __builtin_amdgcn_readfirstlane(threadIdx.x / kTZ);
  const int iz = z0 + lz;
  // loads issued as in the forward tile: no branch waits on one
  const TileList tg = list_begin<LEAN>(a.g_inj.pts, &a.g_inj, a.g_inj.t >= 0, tile, e, x0, z0, g.ld);
  const bool sponge = touches_sponge(g, x0, x0 + TX, z0, z0 + kTZ);
  // (values of points past the grid edge are loaded from element 0 and
  // never used: the update loop skips those points)
  float gp[T::PER], rr[T::PER], dx[T::PER] = {}, dz = 0.f;
  if (sponge) dz = delta_z_load(g, iz);
#pragma unroll
  for (int k = 0; k < T::PER; k++) {
    const int ix = x0 + row * T::PER + k;
    const bool ok = FULL || (ix < g.nx - kHalo && iz < g.nz - kHalo);
    const unsigned q = ok ? (unsigned)(ix * g.ld + iz) : 0u;
    gp[k] = elem(g_prev, q);
    rr[k] = elem(rv, q);
    if (sponge) dx[k] = delta_x_load(g, ix);
  }
  Staged<TX> st;
  stage_load<TX, FULL>(g, a.g_cur + fo, x0, z0, st);
  const PtsLoad li = list_end(tg, &a.g_inj, e);
  stage_store<TX>(st, s_c);
  store_points(li, tp);
  const bool mapped = li.n > 0 && (LEAN || li.n <= kTilePts);
  if (mapped) mark_clear<TX>(map);
  __syncthreads();
  if (mapped) {
    store_amp(li, tp);
    mark_slots(li, tp, map);
  }
  form_u2<P, TX, FULL>(g, s_c, s_u, x0, z0);
  __syncthreads();
  float ucol[T::PER + 2];
  u2_column<TX>(s_u, row, lz, ucol);
#pragma unroll
  for (int k = 0; k < T::PER; k++) {
    const int lx = row * T::PER + k;
    const int ix = x0 + lx;
    if (!FULL && (ix >= g.nx - kHalo || iz >= g.nz - kHalo)) continue;
    const unsigned q = (unsigned)(ix * g.ld + iz);
    const float c = s_c[(lx + kHalo) * T::LZ + lz + kHalo];
    const float* u = s_u + (lx + 1) * T::UZ + lz + 1;
    float gn = sponge ? update_col<true, P>(c, gp[k], rr[k], ucol, k, u, delta_pick(g, iz, dz, dx[k]))
                      : update_col<false, P>(c, gp[k], rr[k], ucol, k, u, 0.f);
    if (mapped) {
      const int h = map[lx * kTZ + lz];
      if (h >= 0) gn = apply_mapped(tp, li.n, h, a.g_inj.sign, q, gn);
    } else if (!LEAN && li.n > 0) {
      gn = apply_points(a.g_inj, tile, e, q, gn);
    }
    elem(g_prev, q) = gn;
  }
}

template <int P, int TX, bool LEAN, bool CHUNKED>
__global__ __launch_bounds__(kThreads) void k_bwd_split(BwdStepArgs a, BlockMap m) {
  using T = Tile<TX>;
  __shared__ float s_c[T::LX * T::LZ];
  __shared__ float s_u[T::UX * T::UZ];
  __shared__ TilePts tp;
  __shared__ __align__(4) signed char map[kTZ * TX];
  int tile, r;
  block_work<CHUNKED>(a.g.ntz * a.g.ntx, m, tile, r);  // r: entry x 2 halves
  const int e = r >> 1;
  const int tx = div_of(tile, m.ntz), model = div_of(e, m.rvdiv);
  const int z0 = kHalo + (tile - tx * a.g.ntz) * kTZ, x0 = kHalo + tx * TX;
  const bool full = tile_is_full<TX>(a.g, x0, z0);
  if (r & 1) {
    if (full)
      bwd_rcv_tile<P, TX, true, LEAN>(a, tile, e, model, x0, z0, s_c, s_u, tp, map);
    else
      bwd_rcv_tile<P, TX, false, LEAN>(a, tile, e, model, x0, z0, s_c, s_u, tp, map);
  } else {
    if (full)
      bwd_src_tile<P, TX, true, LEAN>(a, tile, e, model, x0, z0, s_c, s_u, tp, map);
    else
      bwd_src_tile<P, TX, false, LEAN>(a, tile, e, model, x0, z0, s_c, s_u, tp, map);
  }
}

template <int P, int TX, bool LEAN, bool CHUNKED>
__global__ __launch_bounds__(kThreads) void k_bwd_step(BwdStepArgs a, BlockMap m) {
  using T = Tile<TX>;
  __shared__ float s_sc[T::LX * T::LZ];
  __shared__ float s_gc[T::LX * T::LZ];
  __shared__ float s_su[T::UX * T::UZ];
  __shared__ float s_gu[T::UX * T::UZ];
  __shared__ TilePts s_sub, s_inj;
  __shared__ __align__(4) signed char smap[kTZ * TX];
  __shared__ __align__(4) signed char gmap[kTZ * TX];
  int tile, e;
  block_work<CHUNKED>(a.g.ntz * a.g.ntx, m, tile, e);
  const int tx = div_of(tile, m.ntz), model = div_of(e, m.rvdiv);
  const int z0 = kHalo + (tile - tx * a.g.ntz) * kTZ, x0 = kHalo + tx * TX;
  if (tile_is_full<TX>(a.g, x0, z0))
    bwd_tile<P, TX, true, LEAN>(a, tile, e, model, x0, z0, s_sc, s_gc, s_su, s_gu, s_sub, s_inj, smap, gmap);
  else
    bwd_tile<P, TX, false, LEAN>(a, tile, e, model, x0, z0, s_sc, s_gc, s_su, s_gu, s_sub, s_inj, smap, gmap);
}

__global__ void k_fill(float* p, long long n, float v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_d(double* p, long long n, double v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// One lane per (point, entry).  Points sharing a grid index accumulate in
// the reference's order (manipSource, damp4t10d.cpp:204-214): the first
// active point of each chain applies the whole chain in id order.
__global__ void k_point_inject(float* field, long long fstride, Inject in, int npoints, int nentries) {
  const int pid = blockIdx.x * blockDim.x + threadIdx.x, e = blockIdx.y;
  if (pid >= npoints || e >= nentries) return;
  if (!entry_active(in, e, pid)) return;
  for (int j = in.pts.prev_same[pid]; j >= 0; j = in.pts.prev_same[j])
    if (entry_active(in, e, j)) return;  // an earlier active point leads this chain
  float* f = field + (long long)e * fstride;
  const float* amp = in.amp + e * in.estride + (long long)in.t * in.ld;
  const int q = in.pts.idx[pid];
  float v = f[q];
  for (int j = pid; j >= 0; j = in.pts.next_same[j])
    if (entry_active(in, e, j)) v = in.sign > 0 ? v + amp[j] : v - amp[j];
  f[q] = v;
}

__global__ void k_reciprocal(const float* vel, float* rv, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    rv[i] = 1.0f / vel[i];
}

// A divisor for numerators below `bound` (see Div).
inline Div make_div(int d, long long bound) {
  Div v;
  v.d = d;
  if (d > 1 && bound * d <= (1LL << 32)) v.m = (unsigned)(((1ULL << 32) + (unsigned long long)d - 1) / d);
  return v;
}
// entries: entries of the work list (twice the batch for the split backward,
// whose entries are entry x half); ents: entries of the batch.
inline BlockMap make_map(const Grid& g, int entries, int chunk, int ents, int rvdiv) {
  BlockMap m;
  const int ntiles = g.ntz * g.ntx;
  const long long nblocks = (long long)ntiles * entries;
  m.nentries = entries;
  m.chunk = chunk;
  m.ent = make_div(entries, nblocks);
  m.span = make_div(chunk > 0 ? chunk * entries : 1, nblocks);
  m.chk = make_div(chunk > 0 ? chunk : 1, (long long)chunk * entries);
  m.ntz = make_div(g.ntz, ntiles);
  m.rvdiv = make_div(rvdiv > 0 ? rvdiv : 1, ents);
  return m;
}

inline int grid_for(long long n) {
  long long b = (n + 255) / 256;
  return (int)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
}

template <int TX, bool LEAN, bool CHUNKED>
void fwd_dispatch(const FwdStepArgs& a, int E, bool damped, int precision, hipStream_t s, int chunk) {
  const dim3 grid(a.g.ntz * a.g.ntx * E);
  const BlockMap m = make_map(a.g, E, chunk, E, a.rvdiv);
  if (damped) {
    if (precision == kExact)
      hipLaunchKernelGGL((k_fwd_step<true, kExact, TX, LEAN, CHUNKED>), grid, dim3(kThreads), 0, s, a, m);
    else
      hipLaunchKernelGGL((k_fwd_step<true, kFp32, TX, LEAN, CHUNKED>), grid, dim3(kThreads), 0, s, a, m);
  } else {
    if (precision == kExact)
      hipLaunchKernelGGL((k_fwd_step<false, kExact, TX, LEAN, CHUNKED>), grid, dim3(kThreads), 0, s, a, m);
    else
      hipLaunchKernelGGL((k_fwd_step<false, kFp32, TX, LEAN, CHUNKED>), grid, dim3(kThreads), 0, s, a, m);
  }
}

template <int TX, bool LEAN, bool CHUNKED>
void bwd_dispatch(const BwdStepArgs& a, int E, int precision, hipStream_t s, int chunk) {
  if (a.split) {
    const dim3 grid(a.g.ntz * a.g.ntx * E * 2);
    const BlockMap m = make_map(a.g, 2 * E, chunk, E, a.rvdiv);
    if (precision == kExact)
      hipLaunchKernelGGL((k_bwd_split<kExact, TX, LEAN, CHUNKED>), grid, dim3(kThreads), 0, s, a, m);
    else
      hipLaunchKernelGGL((k_bwd_split<kFp32, TX, LEAN, CHUNKED>), grid, dim3(kThreads), 0, s, a, m);
    return;
  }
  const dim3 grid(a.g.ntz * a.g.ntx * E);
  const BlockMap m = make_map(a.g, E, chunk, E, a.rvdiv);
  if (precision == kExact)
    hipLaunchKernelGGL((k_bwd_step<kExact, TX, LEAN, CHUNKED>), grid, dim3(kThreads), 0, s, a, m);
  else
    hipLaunchKernelGGL((k_bwd_step<kFp32, TX, LEAN, CHUNKED>), grid, dim3(kThreads), 0, s, a, m);
}

// The lean instantiation (no CSR point-list code, no u2 export) serves every
// launch whose active point lists all have slots; the chunked block order
// has lean instantiations only (a general launch walks the
// entries fastest: a different order, the same results).
inline bool fwd_lean(const FwdStepArgs& a) {
  return (a.inj.t < 0 || a.inj.pts.slots) && (a.rec.out == nullptr || a.rec.pts.slots) && a.u2out == nullptr;
}
inline bool bwd_lean(const BwdStepArgs& a) {
  return (a.s_sub.t < 0 || a.s_sub.pts.slots) && (a.g_inj.t < 0 || a.g_inj.pts.slots);
}

template <int TX>
void fwd_dispatch(const FwdStepArgs& a, int E, bool damped, int precision, hipStream_t s, int chunk) {
  if (!fwd_lean(a))
    fwd_dispatch<TX, false, false>(a, E, damped, precision, s, chunk);
  else if (chunk)
    fwd_dispatch<TX, true, true>(a, E, damped, precision, s, chunk);
  else
    fwd_dispatch<TX, true, false>(a, E, damped, precision, s, chunk);
}
template <int TX>
void bwd_dispatch(const BwdStepArgs& a, int E, int precision, hipStream_t s, int chunk) {
  if (!bwd_lean(a))
    bwd_dispatch<TX, false, false>(a, E, precision, s, chunk);
  else if (chunk)
    bwd_dispatch<TX, true, true>(a, E, precision, s, chunk);
  else
    bwd_dispatch<TX, true, false>(a, E, precision, s, chunk);
}

}  // namespace

int tile_count_x(int nx, int tile_x) { return (nx - 2 * kHalo + tile_x - 1) / tile_x; }

void launch_fwd_step(const FwdStepArgs& a, int nentries, bool damped, int precision, hipStream_t stream,
                     int chunk) {
  switch (a.g.tile_x) {
    case 8: fwd_dispatch<8>(a, nentries, damped, precision, stream, chunk); break;
    case 32: fwd_dispatch<32>(a, nentries, damped, precision, stream, chunk); break;
    default: fwd_dispatch<16>(a, nentries, damped, precision, stream, chunk); break;
  }
}

void launch_bwd_step(const BwdStepArgs& a, int nentries, int precision, hipStream_t stream, int chunk) {
  switch (a.g.tile_x) {
    case 8: bwd_dispatch<8>(a, nentries, precision, stream, chunk); break;
    case 32: bwd_dispatch<32>(a, nentries, precision, stream, chunk); break;
    default: bwd_dispatch<16>(a, nentries, precision, stream, chunk); break;
  }
}

void launch_fill(float* p, long long n, float v, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(256), 0, s, p, n, v);
}
void launch_fill_d(double* p, long long n, double v, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_fill_d, dim3(grid_for(n)), dim3(256), 0, s, p, n, v);
}
void launch_point_inject(float* field, long long fstride, const Inject& inj, int npoints, int nentries,
                         hipStream_t s) {
  if (inj.t < 0 || npoints <= 0 || nentries <= 0) return;
  hipLaunchKernelGGL(k_point_inject, dim3((npoints + 255) / 256, nentries), dim3(256), 0, s, field, fstride, inj,
                     npoints, nentries);
}
void launch_reciprocal(const float* vel, float* rv, long long n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_reciprocal, dim3(grid_for(n)), dim3(256), 0, s, vel, rv, n);
}

}  // namespace swfwi
