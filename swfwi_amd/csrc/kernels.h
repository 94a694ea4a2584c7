// Device-side data structures and host launchers of the swfwi HIP kernels.
//
// Layout conventions (identical to the reference, damp4t10d.cpp:138-151,174):
//   * a field is the padded grid nxpad x nzpad, index ix*ld + iz, z contiguous
//     (ld == nzpad);
//   * a batch stores E entries (shots, ensemble members, line-search trials)
//     back to back, entry e at offset e*fstride floats;
//   * the model is kept as its correctly rounded reciprocal rv = 1/vel_t
//     (vel_t = dx^2/(v^2 dt^2), so rv is the squared Courant number).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace swfwi {

constexpr int kHalo = 6;     // stencil reach of one step: 5 (10th order) + 1 (Laplacian of u2)
constexpr int kTZ = 64;      // tile outputs along z (one wavefront row)
constexpr int kTX = 16;      // tile outputs along x
constexpr int kThreads = 256;

enum Precision : int {
  kExact = 0,  // the reference's mixed FP32/FP64 expression trees, bit-for-bit
  kFp32 = 1,   // pure FP32 with fused multiply-adds
};

// Points of interest (sources, receivers) bucketed by stencil tile, so that
// the thread owning a grid point finds the points that touch it with a short
// scan of its tile's bucket.
constexpr int kTilePts = 64;      // points of one tile a block stages in LDS

struct PointList {
  const int* tile_ptr = nullptr;  // [ntiles+1] CSR offsets
  const int* items = nullptr;     // point ids, grouped by tile, ascending id inside a tile
  const int* idx = nullptr;       // [npoints] flattened grid index ix*ld + iz
  // The same lists as fixed-size slots, read with one level of loads: set
  // when no tile holds more than kTilePts points, else null (CSR only).
  const int* count = nullptr;     // [ntiles] points per tile
  // [ntiles][kTilePts] (grid index, point id, position in the tile
  // (x - x0) * 64 + (z - z0), 1 if another point of the tile shares the grid
  // index); grid index -1 past count
  const int4* slots = nullptr;
  // points sharing a grid index, chained in id order (-1 ends a chain)
  const int* prev_same = nullptr; // [npoints]
  const int* next_same = nullptr; // [npoints]
};

struct Grid {
  int nx = 0, nz = 0, ld = 0;     // padded dims and leading dimension
  int ntz = 0, ntx = 0;           // tiles along z and x
  int tile_x = 16;                // tile width along x: 8, 16 or 32
  int nb = 0;                     // damping width handed to the damped kernel (bx0)
  int bx0 = 0, bxn = 0, bz0 = 0, bzn = 0;
  const float* delta_x = nullptr; // damping delta per ix (x sponge), 0 inside
  const float* delta_z = nullptr; // damping delta per iz (bottom sponge)
  int zdamp = 0;                  // first iz of the bottom sponge (nz - nb)
};

// Injection of point amplitudes into a freshly computed field.
struct Inject {
  PointList pts;
  const float* amp = nullptr;     // row-major [t][ld_amp] (+ e*estride per entry)
  int ld = 0;                     // amplitude row length
  long long estride = 0;          // per-entry amplitude offset (0 = shared)
  int t = -1;                     // amplitude row used this step; <0 = no injection
  int sign = 1;                   // +1 addSource, -1 subSource
  const int* lo = nullptr;        // per-entry active id range [lo[e], hi[e]) (nullptr = all)
  const int* hi = nullptr;
};

// Gather of field values at points (receivers) into a trace row.
struct Record {
  PointList pts;
  float* out = nullptr;           // row for this step (+ e*estride)
  long long estride = 0;
};

struct FwdStepArgs {
  Grid g;
  long long fstride = 0;          // floats between entries' fields
  const float* cur = nullptr;     // read-only this step (entry e at + e*fstride)
  float* prev = nullptr;          // overwritten with next
  const float* rv = nullptr;
  long long rvstride = 0;         // 0 = one model shared by all entries
  int rvdiv = 1;                  // entry e uses model e / rvdiv
  Inject inj;                     // applied to next (the writer of the next step's cur)
  Record rec;                     // recorded from cur
  float* band = nullptr;          // boundary band of cur for this step (+ e*band_estride)
  long long band_estride = 0;
  float* u2out = nullptr;         // optional: the u2 scratch of the reference kernel
};

struct BwdStepArgs {
  Grid g;
  long long fstride = 0;
  // source wavefield, reconstructed backwards with the undamped kernel
  const float* s_cur = nullptr;
  float* s_prev = nullptr;
  // receiver wavefield, propagated with the damped kernel
  const float* g_cur = nullptr;
  float* g_prev = nullptr;
  const float* rv = nullptr;
  long long rvstride = 0;
  int rvdiv = 1;
  Inject s_sub;                   // subSource into the new source field (t = it)
  Inject g_inj;                   // receiver injection for the next iteration (t = it-1)
  const float* band = nullptr;    // boundary band restored into the new source field (step it-1)
  const float* band_next = nullptr;  // the band the next launch restores (read ahead into the caches)
  long long band_estride = 0;
  float* grad = nullptr;          // imaging condition accumulator (+ e*gstride)
  long long gstride = 0;
  float weight = 0.f;             // taper weight of this step; xcorr skipped when do_xcorr == 0
  int do_xcorr = 0;
  int split = 0;                  // launch the two fields as separate blocks (small batches)
};

// ---- launchers (all asynchronous on `stream`) -------------------------------
// Tile widths the step kernels are instantiated for, and the tile count along
// x of a padded grid for one of them.
constexpr int kNumWidths = 3;
constexpr int kTileWidths[kNumWidths] = {8, 16, 32};
int tile_count_x(int nx, int tile_x);
// chunk: block order, 0 = entries fastest, else runs of `chunk` tiles per entry
void launch_fwd_step(const FwdStepArgs& a, int nentries, bool damped, int precision, hipStream_t stream,
                     int chunk = 0);
void launch_bwd_step(const BwdStepArgs& a, int nentries, int precision, hipStream_t stream, int chunk = 0);

void launch_fill(float* p, long long n, float v, hipStream_t s);
void launch_fill_d(double* p, long long n, double v, hipStream_t s);
// cur[idx] = cur[idx] (+|-) amp for the active points, sequentially per entry
void launch_point_inject(float* field, long long fstride, const Inject& inj, int npoints, int nentries,
                         hipStream_t s);
void launch_reciprocal(const float* vel, float* rv, long long n, hipStream_t s);

}  // namespace swfwi
