// Mixed-size batched gradients with reduced-precision wavefield checkpoints:
// shots of models on grids of different sizes (and time-step counts) advance
// together, one launch per time step for the whole batch, each entry with
// its own geometry, model, source, receivers and tile range.  The forward
// run stores the source field every kSnapEvery steps in FP32, BF16 or FP16
// with a per-snapshot max-abs scale; the backward run propagates only the
// receiver field and forms the imaging condition from the snapshots instead
// of reconstructing the source field from the saved boundary band
// (damp4t10d.cpp:447-536 inside fwiframework.cpp:114-177).  The stencil
// computes in FP32 (with fused multiply-adds).
#pragma once
#include <cstdint>
#include <functional>
#include <vector>

#include "device.h"

namespace swfwi {

constexpr int kSnapEvery = 4;
// Step-kernel tile: 64 z x 16, 24 or 32 x output points per block of 256
// threads (each thread 4, 6 or 8 consecutive x at one z).  The rule: 24 when
// the batch holds at least kBatchWideTiles tiles of 64 x 16, else 16
// (measured with a model's shots tile by tile: 16 entries of the 501x1501
// grid, 14 112 such tiles, forward + backward 300.6 ms at 24 against 310.4 at
// 32 and 329.4 at 16; the 112 mixed-size entries 1033 ms at 24 against 1068 at
// 32; 16 entries of the 793x277 padded coarse grid, 3 920 tiles, 54.4 ms at
// 16, 55.1 at 24, 61.1 at 32).  At 24 a block's LDS (17.8 KB) leaves room for
// 8 blocks per CU, at 32 (22.4 KB) for 7.
constexpr int kBatchTileZ = 64;
constexpr int kBatchWideTiles = 8000;
constexpr int kBlkTileBits = 20;  // tiles per entry < 2^20, entries < 2^11
constexpr int kRuleOrder = 1;      // block order without tuning: a model's shots tile by tile
constexpr float kBatchTuneMargin = 0.03f;
enum SnapPrec : int { kSnapF32 = 0, kSnapBF16 = 1, kSnapF16 = 2, kSnapNone = 3 };
// Stencils: the reference's damped 10th-order-in-space, 4th-order-in-time
// scheme (fd4t10s-damp-zjh.c), or a damped 4th-order-in-space, 2nd-order-in-
// time leapfrog for coarse grids (next = (2-2d+d^2)p - (1-2d)prev + r Lap4 p).
enum StencilOrder : int { kOrder10 = 10, kOrder4 = 4 };

// One grid size of the batch: the physical model nz x nx with sponge nb,
// nt steps of dt, Ricker sources of peak fm at (sx[s], sz[s]), receivers at
// physical x = gx0 .. gx0+ng-1 on depth gz, nmodels models x nshots shots.
struct BatchGroup {
  int nx = 0, nz = 0, nb = 0;
  float dx = 0, dt = 0, fm = 0;
  int nt = 0;
  std::vector<int> sx, sz;
  int gx0 = 0, ng = 0, gz = 0;
  int nmodels = 1;
  // sponge strength: delta = 0.05 * dscale * dist^2 (1 = the reference's
  // fd4t10s-damp-zjh.c:64-80; a grid stepping k times the reference's dt
  // uses k for the same damping per unit time)
  float dscale = 1.f;
};

// Device description of one entry (one shot of one model).
struct BatchEntry {
  int nxpad, nzpad;      // padded grid; z contiguous (ld = nzpad)
  int ntz, ntx;          // tiles over the update region [6, n-6)
  int damp;              // width of the sponge profile (bx0 = nb + 6)
  float dscale;          // sponge strength factor
  int nt, it_last;       // forward steps; last backward iteration run
  int src;               // index of the source point
  int rec0, ng;          // first receiver point index, count (stride nzpad)
  int nsnap;             // snapshots stored (C[4k+3], k < nsnap)
  int scale0;            // first snapshot scale slot
  int amp0, wts0;        // offsets of the wavelet [nt] and imaging weights [nt]
  long long f;           // field offset (floats) in each of the two field buffers
  long long model;       // offset of the model (1 / vel_t) in the model buffer
  long long snap;        // offset of snapshot 0 (elements)
  long long dcal;        // offset of the [nt][ng] seismogram / adjoint source
  long long grad;        // offset of the entry's gradient (padded grid)
  int tap_nx, tap_nz;    // grid of the 2x-coarser restriction of the field
  int tap_oz;            // fine row of coarse row 0: 2k + tap_oz (physical)
  long long tap;         // offset of the entry's first tap field
};

struct BatchKernelArgs {
  const BatchEntry* ent;
  const int* blk;        // per block of a launch: entry << kBlkTileBits | its tile
  const float* rv;
  const float* amp;
  const float* wts;
  float* fa;             // field buffers: the step reads cur from one, prev/next the other
  float* fb;
  void* snap;
  unsigned* scale;       // per-snapshot max |field| (float bits)
  float* dcal;           // forward: seismogram out; backward: adjoint source in
  float* grad;
  int it_or_k;           // forward: step it; backward: iteration count k (it = nt - 1 - k)
  int rot;               // buffer rotation: even reads cur from fb, odd from fa
  int nblocks;
  float* tap;            // restrictions of cur to the 2x-coarser grid (or null)
  int tap_k;             // tap slot written by this launch, -1 for none
};

void launch_batch_init_fwd(const BatchKernelArgs& a, int nentries, hipStream_t s);
void launch_batch_fwd(const BatchKernelArgs& a, int prec, int order, int tile_x, hipStream_t s);
void launch_batch_init_bwd(const BatchKernelArgs& a, int nentries, hipStream_t s);
void launch_batch_bwd(const BatchKernelArgs& a, int prec, int order, int tile_x, hipStream_t s);
void launch_fill_u32(unsigned* p, long long n, unsigned v, hipStream_t s);
void launch_sub(const float* a, const float* b, float* out, long long n, hipStream_t s);  // out = a - b

class ConvNet;

class Batch {
 public:
  // groups are kept in the given order for models and results; entries run
  // sorted by time-step count (longest first) so the active set of every
  // launch is a prefix of the entry list
  Batch(const std::vector<BatchGroup>& groups, int snap_prec, int order = kOrder10);
  ~Batch();
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;

  int entries() const { return (int)ent_.size(); }
  int models() const { return nmodels_; }
  int snap_prec() const { return prec_; }
  void set_snap_prec(int p);
  void set_graphs(bool on) { graphs_ = on; }
  // Tuner (on by default): the first forward run first times the step
  // kernels (forward and, with checkpoints, backward) over every entry for
  // each free tile width (16, 24, 32) and block order, and keeps the rules'
  // choice unless another is faster by kBatchTuneMargin in both timed
  // rounds.  Off: the rules.
  void set_tuning(bool on);
  // A fixed tile width (16, 24 or 32), or 0 for the rule and the tuner.
  void set_tile_x(int tx);
  int tile_x() const { return tile_x_; }
  // Block order of a launch: 0 entry by entry (the tiles of one entry, then
  // the next), 1 the shots of each model tile by tile (a model tile is read
  // from HBM once and from L2 by the model's other shots), or -1 for the
  // rule (1) and the tuner.
  void set_block_order(int order);
  int block_order() const { return blk_order_; }
  // model m of group g: padded transformed velocity (Damp4t10d layout)
  void bind_model(int g, int m, const float* vel_t);
  // Forward run of every entry: records the seismograms and stores the
  // snapshots.  dcal_host (may be null) receives [entry][nt][ng] in the
  // group / model / shot order of the constructor.
  void forward(float* dcal_host);
  // Backward run: adjoint sources [entry][nt][ng] (same order), raw
  // imaging condition per entry into grad_host [entry][nxpad][nzpad]
  // (may be null); per-model sums over shots into model_grad_host.
  void backward(const float* vsrc_host, float* grad_host, float* model_grad_host);
  // Device adjoint sources from observed data dobs [entry][nt][ng]:
  // vsrc = transVsrc(dobs - dcal) (fwiframework.cpp:75-105,293-302, no mute).
  void adjoint_from_observed(const float* dobs_host, float* vsrc_host);
  // Field taps: the forward run stores the 2x-coarser restriction of the
  // current field at each listed step into tap slots
  // [entry][slot][tap_nx][tap_nz]: physical fine cell (2i, 2k + oz) ->
  // coarse cell (i, k) of the grid with half the nb and (nx+1)/2 x
  // (nz+1-oz)/2 cells (padded: fine (2i-6, 2k-6+oz) -> coarse (i, k)).
  // With oz = 1 both grids have their zero top rows (the free surface) at
  // the same depth.  nb must be even.
  void set_taps(const std::vector<int>& steps, int oz = 0);
  void download_taps(float* host);  // [entry][slot][tap_nx][tap_nz], constructor order
  long long tap_floats() const { return tap_len_; }
  // Run steps it0 .. it0+nsteps-1 of every entry from the given state
  // (prev = field at it0-1, cur = field at it0, [entry][nxpad][nzpad] each,
  // constructor order), injecting the sources as the full run does, and
  // return the state after them (prev = field at it0+nsteps-1, cur = at
  // it0+nsteps).  No recording, checkpoints or taps; the forward hook (the
  // learned correction) runs at the same steps as in forward runs.
  void step_from(const float* prev_host, const float* cur_host, int it0, int nsteps, float* prev_out,
                 float* cur_out);
  // A hook run on the stream after forward step `it` (the fields as
  // (prev, cur) device base pointers, entry stride and count): the learned
  // correction.  Called when (it + 2) % kSnapEvery == 0.
  using StepHook = std::function<void(float* prev, float* cur, long long stride, int entries, hipStream_t s)>;
  void set_forward_hook(StepHook h);
  // The learned correction every kSnapEvery steps of the forward runs
  // (p <- p + net(p) on both time levels, over the physical interior), or
  // off when net is null.  Single-group batches only.
  void set_correction(const ConvNet* net);
  int order() const { return order_; }
  hipStream_t stream() const { return stream_.get(); }
  // point updates of one forward and one backward run, and their launch counts
  double forward_points() const;
  double backward_points() const;
  long long snapshot_bytes() const { return snap_bytes_; }
  double last_ms[2] = {0, 0};  // forward, backward wall time on the stream

  struct Layout {  // host copy of an entry's placement, constructor order
    int group, model, shot;
    int nxpad, nzpad, nt, ng;
    long long dcal, grad;
  };
  const std::vector<Layout>& layout() const { return lay_; }

 private:
  void build();
  void layout_tiles(int tx, int order);  // entry tile counts and the block list
  void tune();
  void drop_graphs();
  void run_loop(bool fwd);
  BatchKernelArgs args(bool fwd);
  std::vector<BatchGroup> g_;
  int prec_;
  int order_;
  std::vector<int> taps_;
  long long tap_len_ = 0;  // floats of every entry's tap slots
  DBuf<float> tap_;
  StepHook hook_;
  DBuf<float> ycorr_;
  bool graphs_ = true;
  int nmodels_ = 0;
  std::vector<BatchEntry> ent_;      // run order (nt descending)
  std::vector<int> run_of_;          // constructor order -> run index
  std::vector<Layout> lay_;
  std::vector<int> tile0_;           // per run index, prefix of tiles
  int tile_x_ = 16;                  // step-kernel tile width
  int rule_x_ = 16;                  // the size rule's width
  bool tune_ = true, tuned_ = false;
  int fixed_x_ = 0;                  // a width set by the caller, 0 for none
  int blk_order_ = 1;                // block order in use
  int fixed_order_ = -1;             // an order set by the caller, -1 for none
  std::vector<long long> model_off_; // per model (groups in order)
  std::vector<int> model_group_;
  long long fields_ = 0, models_len_ = 0, dcal_len_ = 0, grad_len_ = 0, snap_elems_ = 0;
  int scales_ = 0;
  long long snap_bytes_ = 0;
  Stream stream_;
  DBuf<BatchEntry> dent_;
  DBuf<int> dblk_;
  DBuf<float> rv_, amp_, wts_, fa_, fb_, dcal_, vsrc_, grad_, mgrad_;
  DBuf<unsigned char> snap_;
  DBuf<unsigned> scale_;
  hipGraphExec_t graph_[2] = {nullptr, nullptr};
};

}  // namespace swfwi
