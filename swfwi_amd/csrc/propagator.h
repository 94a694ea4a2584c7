// The device-resident acoustic propagator: the MI355X counterpart of
// Damp4t10d's time loops (damp4t10d.cpp:272-307) and of calgradient
// (fwiframework.cpp:114-177, essfwiframework.cpp:114-173), batched over
// entries (shots, ensemble members, line-search trials).
#pragma once
#include <cstdint>
#include <functional>
#include <memory>
#include <vector>

#include "device.h"
#include "kernels.h"
#include "ops.h"

namespace swfwi {

// Acquisition + grid description, in the reference's terms: physical model
// size nz x nx, sponge nb, positions as physical cell indices
// (ShotPosition::getx/getz, shot-position.cpp:11-37).
struct Geometry {
  int nx = 0, nz = 0, nb = 0;
  float dt = 0, dx = 0, fm = 0;
  int nt = 0;
  std::vector<int> sx, sz, gx, gz;

  // derived (Damp4t10d ctor + expandDomain, damp4t10d.cpp:138-151,367-376)
  int nxpad = 0, nzpad = 0, ld = 0;
  int bx0 = 0, bxn = 0, bz0 = 0, bzn = 0;
  int bsize = 0;      // floats of one step's boundary band (initBndryVector)
  int ntz = 0, ntx = 0;
  int ns() const { return (int)sx.size(); }
  int ng() const { return (int)gx.size(); }
  long long npts() const { return (long long)nxpad * ld; }
  void derive();      // fill the derived members; throws kErrArg on bad input
  GridGeom gridgeom() const { return GridGeom{nxpad, nzpad, ld, bx0, bxn, bz0, bzn}; }
};

// Per-tile CSR of a point set on the device.
struct DevPoints {
  DBuf<int> tile_ptr, items, idx, count, prev_same, next_same;
  DBuf<int4> slots;
  int n = 0;
  int max_per_tile = 0;
  PointList view() const {
    PointList p;
    p.tile_ptr = tile_ptr.get();
    p.items = items.get();
    p.idx = idx.get();
    p.prev_same = prev_same.get();
    p.next_same = next_same.get();
    if (max_per_tile <= kTilePts && slots.get()) {
      p.count = count.get();
      p.slots = slots.get();
    }
    return p;
  }
};

// Damping tables of the damped kernel for a padded grid and width nb.
struct DampTables {
  DBuf<float> dx, dz;
  int zdamp = 0;
};
void build_damping(int nx, int nz, int nb, DampTables& t, hipStream_t s);
// Bucket point indices (ix*ld + iz) by the tiles of width tile_x whose
// output region, grown by `halo` cells, contains them.
void build_points(const std::vector<long long>& idx, int ld, int ntz, int ntx, int tile_x, int halo, DevPoints& out,
                  hipStream_t s);

// Wavefield workspace for E entries (two buffers per wavefield).
struct Fields {
  int E = 0;
  DBuf<float> a, b, c, d;       // four field buffers per entry
  float* s_cur = nullptr;       // source field state after forward(): cur of the last step
  float* s_nxt = nullptr;       //   ... and the field it produced
  float* spare0 = nullptr;      // the two buffers forward() left free (the
  float* spare1 = nullptr;      //   receiver field pair of backward())
  void ensure(int E, long long npts, bool receiver);
};

class Propagator {
 public:
  Propagator(const Geometry& g, int precision);

  const Geometry& geo() const { return g_; }
  int precision() const { return prec_; }
  void set_precision(int p) { prec_ = p; }
  hipStream_t stream() const { return stream_.get(); }
  void sync() const { stream_.sync(); }
  int device() const { return dev_; }  // the HIP device current at creation
  Grid grid(int w = 1) const;  // w indexes kTileWidths
  const int* src_xp() const { return sxp_.get(); }  // padded coordinates (device)
  const int* src_zp() const { return szp_.get(); }
  const int* rec_xp() const { return gxp_.get(); }
  const int* rec_zp() const { return gzp_.get(); }

  // Source amplitudes for forward()/backward(): row t of entry e is
  // amp + e*estride + t*ld; entry e injects source ids [lo[e], hi[e]) (device
  // arrays; nullptr = every source).
  struct Sources {
    const float* amp = nullptr;
    int ld = 0;
    long long estride = 0;
    const int* lo = nullptr;
    const int* hi = nullptr;
  };

  // nt steps of the damped forward propagation for E entries with models
  // rv + e*rvstride.  Records [E][nt][ng] into dcal and/or saves the boundary
  // band [E][nt][bsize] when those are non-null.
  void forward(int E, const float* rv, long long rvstride, const Sources& src, float* dcal, float* band, Fields& w,
               int rvdiv = 1);

  // One forward run of one entry (the Sources / dcal contract of forward())
  // over two caller-provided field buffers of npts() floats each, on the
  // handle's stream: no graph, no timing, no workspace growth and the fixed
  // tile rule instead of the tuner, so nothing is allocated.  No band.
  void forward_external(const float* rv, const Sources& src, float* dcal, float* fa, float* fb);

  // The backward half of calgradient on the state forward() left in w:
  // vsrc [E][nt][ng] adjoint sources (+e*vstride), grad accumulates the
  // imaging condition (+e*gstride; caller zeroes it).  Returns the number of
  // backward iterations run.
  int backward(int E, const float* rv, long long rvstride, const Sources& src, const float* vsrc, long long vstride,
               const float* band, float* grad, long long gstride, Fields& w, int rvdiv = 1);

  // The backward iterations calgradient performs with an imaging condition
  // (it = nt-1 down to the returned value), with their taper weights.
  static int xcorr_range(float dt, int nt, std::vector<float>& weights);

  // launch counters for throughput accounting
  long long fwd_launches = 0, bwd_launches = 0;
  double fwd_point_updates = 0, bwd_point_updates = 0;

  // Time loops are captured into HIP graphs and replayed when the same
  // buffers come back (the drivers' epochs); off by default for one-shot calls.
  // Mode 1 captures batches of up to kGraphMaxEntries entries (larger ones
  // are not launch-bound: measured on the 301x921 grid at 100 members, an
  // ensemble FWI epoch ran 2585 ms launched directly against 2612 ms
  // replayed), mode 2 captures every batch size.
  void set_graphs(int mode) { graphs_ = mode; }
  // Tile width override (8, 16, 32) or 0 for the tuner; tuning off falls back
  // to the measured rule by batch size.
  void set_tile_x(int tx);
  // The tile tuner (on by default): the first forward() of a batch size
  // times every width on scratch wavefields and keeps the measured rule's
  // width unless another is faster by kTuneMargin in both timed rounds.
  void set_tuning(bool on) {
    tune_ = on;
    tile_cache_.clear();
  }
  bool tuning() const { return tune_; }
  static constexpr float kTuneMargin = 0.03f;
  // Backward iterations as one fused block per tile (0), as separate source
  // and receiver blocks (1), or by batch size (-1, the default: split while
  // the fused tiles would not fill the chip).
  void set_bwd_split(int mode);
  bool bwd_split(int E) const;
  int tile_choice_cached(int kind, int E) const;
  long long tile_key(int kind, int E) const;
  // block order of a step launch (kind 0 forward, 1 backward): 0 = entries
  // fastest, else runs of that many tiles per entry
  int block_chunk(int kind, int E, long long rvstride, int rvdiv) const;
  // Microseconds per launch of `reps` back-to-back step kernels (kind 0
  // forward, 1 backward as backward() launches it) of tile width kTileWidths[width] on E scratch
  // wavefields of model rv (entry e uses rv + (e / rvdiv) * rvstride).
  float time_steps(int kind, int E, const float* rv, long long rvstride, int rvdiv, Fields& w, int width, int reps,
                   bool clear);
  bool graphs() const { return graphs_ != 0; }
  int graphs_mode() const { return graphs_; }
  // Timing: one HIP event pair around every forward (0) / backward (1) time
  // loop on the handle's stream, graph replays included; collect_timing()
  // waits for the loops in flight and adds them to kernel_ms / kernel_n
  // (milliseconds, step launches).
  void set_timing(bool on);
  void collect_timing();
  // pending timed loops collected (host waits) once this many are in flight
  void set_timing_flush(int n) { timing_flush_ = n < 1 ? 1 : n; }
  double kernel_ms[2] = {0, 0};
  long long kernel_n[2] = {0, 0};

  ~Propagator();
  Propagator(const Propagator&) = delete;
  Propagator& operator=(const Propagator&) = delete;

 private:
  // one launch of a forward time loop: step `it` writes next over prev
  struct FwdLaunch {
    int it;
    float *cur, *prev;
  };
  // the launches of a forward run over fw's buffers; sets fw's final state
  std::vector<FwdLaunch> plan_forward(Fields& fw) const;
  // one forward run of E entries on stream s: p(-1) in fa, p(0) in fb, the
  // two buffers then alternate (the rotation plan_forward describes)
  void forward_body(int E, const float* rv, long long rvstride, const Sources& src, float* dcal, float* band,
                    float* fa, float* fb, int rvdiv, int w, hipStream_t s);
  void backward_body(int E, const float* rv, long long rvstride, const Sources& src, const float* vsrc,
                     long long vstride, const float* band, float* grad, long long gstride, Fields& fw, int rvdiv,
                     int it_min, const std::vector<float>& wts, int w);
  // Tile width (index into kTileWidths) for a kernel kind (0 forward,
  // 1 backward) and batch size: fixed override, cached choice, or a timed
  // trial of every width on scratch wavefields (the tile tuner).
  int tile_choice(int kind, int E, const float* rv, long long rvstride, int rvdiv, Fields& w, bool may_time);
  // the measured width rule (index into kTileWidths) by batch size, or the override
  int fixed_width(int kind, int E) const;
  static constexpr int kGraphMaxEntries = 8;
  // A forward batch of 2..kLanes entries without a band (line-search trial
  // runs) runs as one-entry loops side by side, lane e on its own stream:
  // measured on the 301x921 grid, two one-entry loops on two streams take
  // 25 ms where one two-entry loop takes 30 ms.
  static constexpr int kLanes = 4;
  static constexpr long long kBwdSplitBelow = 6000;  // 64x16 tiles over the batch
  // Run a time loop (`launches` step launches of kind 0 forward / 1
  // backward): captured into a HIP graph on first use and replayed when the
  // same buffers come back, or launched directly.
  void run_maybe_graph(const std::vector<uintptr_t>& key, int E, int kind, long long launches,
                       const std::function<void()>& body, hipStream_t s);
  hipEvent_t pooled_event();
  void flush_timing_if_full();
  int timing_flush_ = 64;

  struct GraphEntry {
    std::vector<uintptr_t> key;
    hipGraphExec_t exec = nullptr;
  };
  struct TimedLoop {
    hipEvent_t start, stop;
    int kind;
    long long launches;
  };
  std::vector<GraphEntry> graph_cache_;
  int graphs_ = 0;
  bool timing_ = false;
  std::vector<TimedLoop> pending_;
  std::vector<hipEvent_t> event_pool_;

  Geometry g_;
  int prec_;
  int dev_ = 0;
  Stream stream_;
  Stream lane_streams_[kLanes - 1];  // lanes 1.. (lane 0 is stream_)
  hipEvent_t lane_events_[kLanes] = {};
  DampTables damp_;
  DevPoints src_[kNumWidths], rec_[kNumWidths];  // points per tile, per tile width
  int tile_override_ = 0;                 // 0 = tune, else a width from kTileWidths
  bool tune_ = true;
  DBuf<float> tune_grad_;  // scratch gradient of the tuner's backward trials
  int bwd_split_ = -1;
  std::vector<std::pair<long long, int>> tile_cache_;  // (kind*2^32 + E) -> width index
  DBuf<int> sxp_, szp_, gxp_, gzp_;
};

}  // namespace swfwi
