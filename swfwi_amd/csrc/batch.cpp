// Host side of the mixed-size batch (batch.h): entry tables, buffers and
// the forward / backward time loops, captured into HIP graphs.
#include "batch.h"

#include <algorithm>
#include <cmath>
#include <functional>
#include <numeric>
#include <string>

#include "hostmath.h"
#include "kernels.h"
#include "learned.h"
#include "ops.h"

namespace swfwi {

namespace {

long long round64(long long n) { return (n + 63) / 64 * 64; }

// the imaging weight of backward iteration `it` (fwiframework.cpp:164-175:
// float dt*it against double thresholds), 0 where the reference breaks
float taper(float dt, int it) {
  if (dt * it > 0.4) return 1.f;
  if (dt * it > 0.3) return (float)((dt * it - 0.3) / 0.1);
  return 0.f;
}

}  // namespace

Batch::Batch(const std::vector<BatchGroup>& groups, int snap_prec, int order)
    : g_(groups), prec_(snap_prec), order_(order) {
  if (g_.empty()) fail(kErrArg, "a batch needs at least one group");
  if (prec_ != kSnapF32 && prec_ != kSnapBF16 && prec_ != kSnapF16 && prec_ != kSnapNone)
    fail(kErrArg, "unknown snapshot precision");
  if (order_ != kOrder10 && order_ != kOrder4) fail(kErrArg, "unknown stencil order");
  build();
}

Batch::~Batch() {
  for (hipGraphExec_t& g : graph_)
    if (g) (void)hipGraphExecDestroy(g);
}

void Batch::drop_graphs() {
  for (hipGraphExec_t& g : graph_)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

void Batch::set_tuning(bool on) {
  tune_ = on;
  tuned_ = false;
  const int tx = fixed_x_ ? fixed_x_ : rule_x_, order = fixed_order_ >= 0 ? fixed_order_ : kRuleOrder;
  if (tile_x_ != tx || blk_order_ != order) layout_tiles(tx, order);
}

void Batch::set_tile_x(int tx) {
  if (tx != 0 && tx != 16 && tx != 24 && tx != 32)
    fail(kErrArg, "batch tile width must be 0 (rule and tuner), 16, 24 or 32");
  fixed_x_ = tx;
  set_tuning(tune_);
}

void Batch::set_block_order(int order) {
  if (order < -1 || order > 1) fail(kErrArg, "batch block order must be -1 (rule and tuner), 0 or 1");
  fixed_order_ = order;
  set_tuning(tune_);
}

void Batch::layout_tiles(int tx, int order) {
  hipStream_t s = stream_.get();
  SWFWI_HIP(hipStreamSynchronize(s));  // no launch in flight reads the old layout
  tile_x_ = tx;
  blk_order_ = order;
  // entries run longest first and a model's shots are consecutive entries of
  // one grid and one step count, so every launch's active entries (a prefix)
  // hold whole runs of shots of a model: each run's blocks are contiguous
  std::vector<int> blk;
  tile0_.assign(ent_.size() + 1, 0);
  for (BatchEntry& e : ent_) e.ntx = (e.nxpad - 2 * kHalo + tx - 1) / tx;
  for (size_t r = 0; r < ent_.size();) {
    size_t r1 = r + 1;
    while (r1 < ent_.size() && ent_[r1].model == ent_[r].model) r1++;
    const int nt_e = ent_[r].ntz * ent_[r].ntx;
    if (nt_e >= (1 << kBlkTileBits)) fail(kErrRange, "too many tiles per batch entry");
    if (order == 1) {
      for (int t = 0; t < nt_e; t++)
        for (size_t q = r; q < r1; q++) blk.push_back((int)(q << kBlkTileBits) | t);
    } else {
      for (size_t q = r; q < r1; q++)
        for (int t = 0; t < nt_e; t++) blk.push_back((int)(q << kBlkTileBits) | t);
    }
    for (size_t q = r; q < r1; q++) tile0_[q + 1] = tile0_[q] + nt_e;
    r = r1;
  }
  dent_.upload(ent_.data(), ent_.size(), s);
  dblk_.upload(blk, s);
  SWFWI_HIP(hipStreamSynchronize(s));
  drop_graphs();
}

void Batch::tune() {
  // Timed trial on the batch's own buffers before its first forward run,
  // which clears or overwrites everything the trial writes (fields, scales
  // and taps are cleared; seismogram rows and checkpoint slots are written
  // step by step; the backward run clears the gradient).  kReps steps from
  // step 0 (forward) and iteration 0 (backward), where every entry is
  // active; three interleaved rounds, the first warming clocks and caches.
  tuned_ = true;
  constexpr int kReps = 8;
  const int n = (int)ent_.size();
  int min_fwd = ent_[0].nt, min_bwd = ent_[0].nt;
  for (const BatchEntry& e : ent_) {
    min_fwd = std::min(min_fwd, e.nt);
    min_bwd = std::min(min_bwd, e.nt - e.it_last);
  }
  if (min_fwd < kReps) return;
  const bool bwd = prec_ != kSnapNone && min_bwd >= kReps;
  hipStream_t s = stream_.get();
  launch_fill(fa_.get(), fields_, 0.f, s);
  launch_fill(fb_.get(), fields_, 0.f, s);
  launch_fill_u32(scale_.get(), std::max(scales_, 1), 0u, s);
  if (bwd) {
    launch_fill(vsrc_.get(), dcal_len_, 0.f, s);
    launch_fill(grad_.get(), grad_len_, 0.f, s);
    if (snap_bytes_ >= 4) launch_fill_u32(reinterpret_cast<unsigned*>(snap_.get()), snap_bytes_ / 4, 0u, s);
  }
  // candidates: the free widths x the free block orders
  std::vector<std::pair<int, int>> cand;
  for (int w : {16, 24, 32})
    for (int o : {0, 1})
      if ((!fixed_x_ || w == fixed_x_) && (fixed_order_ < 0 || o == fixed_order_)) cand.push_back({w, o});
  const std::pair<int, int> rule{fixed_x_ ? fixed_x_ : rule_x_, fixed_order_ >= 0 ? fixed_order_ : kRuleOrder};
  if (cand.size() < 2) return;
  std::vector<float> t[2];
  t[0].assign(cand.size(), 0.f);
  t[1].assign(cand.size(), 0.f);
  hipEvent_t e0, e1;
  SWFWI_HIP(hipEventCreate(&e0));
  SWFWI_HIP(hipEventCreate(&e1));
  for (int round = 0; round < 3; round++)
    for (size_t wi = 0; wi < cand.size(); wi++) {
      layout_tiles(cand[wi].first, cand[wi].second);
      BatchKernelArgs a = args(true);
      a.tap = nullptr;
      SWFWI_HIP(hipEventRecord(e0, s));
      for (int r = 0; r < kReps; r++) {
        a.it_or_k = r;
        a.rot = r;
        a.nblocks = tile0_[n];
        launch_batch_fwd(a, prec_, order_, tile_x_, s);
      }
      if (bwd) {
        BatchKernelArgs b = args(false);
        for (int r = 0; r < kReps; r++) {
          b.it_or_k = r;
          b.rot = r;
          b.nblocks = tile0_[n];
          launch_batch_bwd(b, prec_, order_, tile_x_, s);
        }
      }
      SWFWI_HIP(hipEventRecord(e1, s));
      SWFWI_HIP(hipEventSynchronize(e1));
      float ms = 0;
      SWFWI_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (round > 0) t[round - 1][wi] = ms;
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  SWFWI_HIP(hipGetLastError());
  size_t ri = 0;
  for (size_t i = 0; i < cand.size(); i++)
    if (cand[i] == rule) ri = i;
  size_t best = ri;
  for (size_t i = 0; i < cand.size(); i++) {
    const bool faster = t[0][i] < (1.f - kBatchTuneMargin) * t[0][ri] && t[1][i] < (1.f - kBatchTuneMargin) * t[1][ri];
    if (faster && std::max(t[0][i], t[1][i]) < std::max(t[0][best], t[1][best])) best = i;
  }
  layout_tiles(cand[best].first, cand[best].second);
}

void Batch::set_snap_prec(int p) {
  if (p != kSnapF32 && p != kSnapBF16 && p != kSnapF16 && p != kSnapNone) fail(kErrArg, "unknown snapshot precision");
  if (p == prec_) return;
  // the checkpoint buffer is sized at construction: FP32 holds any format,
  // BF16/FP16 hold each other, a batch without checkpoints holds none
  if (prec_ == kSnapNone || (p == kSnapF32 && prec_ != kSnapF32))
    fail(kErrArg, "the batch's checkpoint storage was created for a narrower format");
  prec_ = p;
  drop_graphs();
  tuned_ = false;  // the kernels change with the checkpoint format
}

void Batch::build() {
  hipStream_t s = stream_.get();
  struct Pending {
    BatchEntry e;
    int ctor;
  };
  std::vector<Pending> list;
  std::vector<float> amp, wts;
  long long model_len = 0, dcal_len = 0, grad_len = 0;
  int mbase = 0;
  // tile width from the batch's size in 64 x 16 tiles
  long long t16 = 0;
  for (const BatchGroup& G : g_) {
    const long long ux = G.nx + 2 * G.nb, uz = G.nz + G.nb;  // update region
    t16 += (long long)G.nmodels * (long long)G.sx.size() * ((uz + kBatchTileZ - 1) / kBatchTileZ) * ((ux + 15) / 16);
  }
  tile_x_ = rule_x_ = t16 >= kBatchWideTiles ? 24 : 16;
  for (size_t gi = 0; gi < g_.size(); gi++) {
    const BatchGroup& G = g_[gi];
    const int ns = (int)G.sx.size();
    if (G.nx < 1 || G.nz < 1 || G.nb < 6 || G.nt < kSnapEvery || ns < 1 || G.ng < 1 || G.nmodels < 1 ||
        G.sz.size() != G.sx.size() || !(G.dt > 0) || !(G.dx > 0) || !(G.dscale > 0))
      fail(kErrArg, "invalid batch group " + std::to_string(gi));
    const int nxpad = G.nx + 2 * G.nb + 2 * kHalo, nzpad = G.nz + G.nb + 2 * kHalo;
    const long long npts = (long long)nxpad * nzpad;
    if (npts >= (1LL << 30)) fail(kErrRange, "grid too large");
    const int bx0 = G.nb + kHalo, bz0 = kHalo;
    if (G.gx0 < 0 || G.gx0 + G.ng > G.nx || G.gz < 0 || G.gz >= G.nz) fail(kErrArg, "receivers outside the model");
    for (int k = 0; k < ns; k++)
      if (G.sx[k] < 0 || G.sx[k] >= G.nx || G.sz[k] < 0 || G.sz[k] >= G.nz) fail(kErrArg, "source outside the model");
    // imaging iterations: it = 4k with a positive taper weight
    int it_last = -1;
    const int amp0 = (int)amp.size(), wts0 = (int)wts.size();
    std::vector<float> w(G.nt, 0.f);
    for (int it = G.nt - 1; it >= kSnapEvery; it--) {
      const float t = taper(G.dt, it);
      if (t == 0.f) break;
      if (it % kSnapEvery == 0) {
        w[it] = (float)kSnapEvery * t;
        it_last = it;
      }
    }
    if (it_last < 0) fail(kErrArg, "nt too short for an imaging condition");
    std::vector<float> wl(G.nt);
    ricker_wavelet(wl.data(), G.nt, G.fm, G.dt, 1.0f);
    amp.insert(amp.end(), wl.begin(), wl.end());
    wts.insert(wts.end(), w.begin(), w.end());
    // the step kernel forms per-point byte offsets in 32 bits
    if (npts >= (1LL << 29)) fail(kErrRange, "padded grid too large for a batch entry");
    for (int m = 0; m < G.nmodels; m++) {
      model_off_.push_back(model_len);
      model_group_.push_back((int)gi);
      model_len += round64(npts);
      for (int k = 0; k < ns; k++) {
        BatchEntry e{};
        e.nxpad = nxpad;
        e.nzpad = nzpad;
        e.ntz = (nzpad - 2 * kHalo + kBatchTileZ - 1) / kBatchTileZ;
        e.ntx = (nxpad - 2 * kHalo + tile_x_ - 1) / tile_x_;
        e.damp = bx0;
        e.dscale = G.dscale;
        e.nt = G.nt;
        e.it_last = it_last;
        e.src = (G.sx[k] + bx0) * nzpad + G.sz[k] + bz0;
        e.rec0 = (G.gx0 + bx0) * nzpad + G.gz + bz0;
        e.ng = G.ng;
        e.nsnap = prec_ == kSnapNone ? 0 : G.nt / kSnapEvery;
        e.tap_nx = (G.nx + 1) / 2 + 2 * (G.nb / 2) + 2 * kHalo;
        e.tap_nz = (G.nz + 1) / 2 + G.nb / 2 + 2 * kHalo;
        e.amp0 = amp0;
        e.wts0 = wts0;
        e.model = model_off_.back();
        e.dcal = dcal_len;
        e.grad = grad_len;
        dcal_len += round64((long long)G.nt * G.ng);
        grad_len += round64(npts);
        lay_.push_back(Layout{(int)gi, m, k, nxpad, nzpad, G.nt, G.ng, e.dcal, e.grad});
        list.push_back(Pending{e, (int)list.size()});
      }
    }
    mbase += G.nmodels;
  }
  nmodels_ = mbase;
  // run order: longest first, so the active entries of every launch are a
  // prefix (the backward length nt - it_last must follow the same order)
  std::stable_sort(list.begin(), list.end(), [](const Pending& a, const Pending& b) { return a.e.nt > b.e.nt; });
  for (size_t i = 1; i < list.size(); i++)
    if (list[i].e.nt - list[i].e.it_last > list[i - 1].e.nt - list[i - 1].e.it_last)
      fail(kErrArg, "entries with fewer steps must not need more backward iterations");
  run_of_.assign(list.size(), 0);
  if (list.size() >= (1u << (31 - kBlkTileBits))) fail(kErrRange, "too many batch entries");
  long long f = 0, snap = 0;
  int scale = 0;
  for (size_t r = 0; r < list.size(); r++) {
    BatchEntry& e = list[r].e;
    const long long npts = (long long)e.nxpad * e.nzpad;
    e.f = f;
    f += round64(npts);
    e.snap = snap;
    snap += (long long)e.nsnap * npts;
    e.scale0 = scale;
    scale += e.nsnap;
    run_of_[list[r].ctor] = (int)r;
    ent_.push_back(e);
  }
  fields_ = f;
  models_len_ = model_len;
  dcal_len_ = dcal_len;
  grad_len_ = grad_len;
  snap_elems_ = snap;
  scales_ = scale;
  const size_t esz = prec_ == kSnapF32 ? 4 : (prec_ == kSnapNone ? 0 : 2);
  snap_bytes_ = (long long)(snap * esz);

  amp_.upload(amp, s);
  wts_.upload(wts, s);
  rv_.resize((size_t)models_len_);
  launch_fill(rv_.get(), models_len_, 0.f, s);
  fa_.resize((size_t)fields_);
  fb_.resize((size_t)fields_);
  dcal_.resize((size_t)dcal_len_);
  vsrc_.resize((size_t)dcal_len_);
  grad_.resize((size_t)grad_len_);
  mgrad_.resize((size_t)models_len_);
  snap_.resize((size_t)std::max(snap_bytes_, 1LL));
  scale_.resize((size_t)std::max(scales_, 1));
  layout_tiles(tile_x_, kRuleOrder);  // uploads the entries and the block list
}

void Batch::bind_model(int g, int m, const float* vel_t) {
  if (g < 0 || g >= (int)g_.size() || m < 0 || m >= g_[g].nmodels) fail(kErrArg, "no such model");
  int mi = m;
  for (int gi = 0; gi < g; gi++) mi += g_[gi].nmodels;
  const BatchGroup& G = g_[g];
  const long long npts = (long long)(G.nx + 2 * G.nb + 2 * kHalo) * (G.nz + G.nb + 2 * kHalo);
  hipStream_t s = stream_.get();
  float* dst = rv_.get() + model_off_[mi];
  SWFWI_HIP(hipMemcpyAsync(dst, vel_t, sizeof(float) * npts, hipMemcpyHostToDevice, s));
  op_reciprocal(dst, dst, npts, s);  // the kernels read 1 / vel_t
  SWFWI_HIP(hipStreamSynchronize(s));
}

BatchKernelArgs Batch::args(bool fwd) {
  BatchKernelArgs a{};
  a.ent = dent_.get();
  a.blk = dblk_.get();
  a.rv = rv_.get();
  a.amp = amp_.get();
  a.wts = wts_.get();
  a.fa = fa_.get();
  a.fb = fb_.get();
  a.snap = snap_.get();
  a.scale = scale_.get();
  a.dcal = fwd ? dcal_.get() : vsrc_.get();
  a.grad = grad_.get();
  a.tap = tap_.get();
  a.tap_k = -1;
  return a;
}

void Batch::run_loop(bool fwd) {
  hipStream_t s = stream_.get();
  BatchKernelArgs a = args(fwd);
  const int n = (int)ent_.size();
  launch_fill(fa_.get(), fields_, 0.f, s);
  launch_fill(fb_.get(), fields_, 0.f, s);
  if (fwd) {
    launch_fill_u32(scale_.get(), std::max(scales_, 1), 0u, s);
    if (tap_len_ > 0) launch_fill(tap_.get(), tap_len_, 0.f, s);
    launch_batch_init_fwd(a, n, s);
    const int steps = ent_[0].nt;
    for (int it = 0; it < steps; it++) {
      int act = n;
      while (act > 0 && ent_[act - 1].nt <= it) act--;
      a.it_or_k = it;
      a.rot = it;
      a.nblocks = tile0_[act];
      a.tap_k = -1;
      for (size_t k = 0; k < taps_.size(); k++)
        if (taps_[k] == it) a.tap_k = (int)k;
      launch_batch_fwd(a, prec_, order_, tile_x_, s);
      if (hook_ && (it + 2) % kSnapEvery == 0) {
        // after step it the buffer step it wrote holds the field at it+1
        float* cur = (it & 1) == 0 ? fa_.get() : fb_.get();
        float* prev = (it & 1) == 0 ? fb_.get() : fa_.get();
        hook_(prev, cur, ent_.size() > 1 ? ent_[1].f - ent_[0].f : 0, act, s);
      }
    }
  } else {
    if (prec_ == kSnapNone) fail(kErrState, "a batch without checkpoints has no backward run");
    launch_fill(grad_.get(), grad_len_, 0.f, s);
    launch_batch_init_bwd(a, n, s);
    const int iters = ent_[0].nt - ent_[0].it_last;
    for (int k = 0; k < iters; k++) {
      int act = n;
      while (act > 0 && ent_[act - 1].nt - ent_[act - 1].it_last <= k) act--;
      a.it_or_k = k;
      a.rot = k;
      a.nblocks = tile0_[act];
      launch_batch_bwd(a, prec_, order_, tile_x_, s);
    }
  }
}

void Batch::set_taps(const std::vector<int>& steps, int oz) {
  for (const BatchGroup& G : g_)
    if (G.nb % 2) fail(kErrArg, "field taps need an even nb (the coarse grid has nb/2)");
  if (oz != 0 && oz != 1) fail(kErrArg, "tap row offset must be 0 or 1");
  for (int t : steps)
    if (t < 0) fail(kErrArg, "negative tap step");
  taps_ = steps;
  long long off = 0;
  for (size_t c = 0; c < lay_.size(); c++) {  // constructor order
    BatchEntry& e = ent_[run_of_[c]];
    const BatchGroup& G = g_[lay_[c].group];
    e.tap_oz = oz;
    e.tap_nz = (G.nz + 1 - oz) / 2 + G.nb / 2 + 2 * kHalo;
    e.tap = off;
    off += (long long)steps.size() * e.tap_nx * e.tap_nz;
  }
  tap_len_ = off;
  tap_.resize((size_t)std::max(tap_len_, 1LL));
  dent_.upload(ent_.data(), ent_.size(), stream_.get());
  SWFWI_HIP(hipStreamSynchronize(stream_.get()));
  for (hipGraphExec_t& g : graph_)  // the captured loops bake the tap slots in
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

void Batch::download_taps(float* host) {
  if (tap_len_ > 0)
    SWFWI_HIP(hipMemcpy(host, tap_.get(), sizeof(float) * tap_len_, hipMemcpyDeviceToHost));
}

void Batch::set_forward_hook(StepHook h) {
  if (h && g_.size() != 1) fail(kErrArg, "a forward hook needs a single-group batch");
  hook_ = std::move(h);
  for (hipGraphExec_t& g : graph_)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

void Batch::set_correction(const ConvNet* net) {
  if (!net) {
    set_forward_hook(nullptr);
    return;
  }
  if (g_.size() != 1) fail(kErrArg, "the learned correction needs a single-group batch");
  const BatchGroup& G = g_[0];
  const int nxpad = G.nx + 2 * G.nb + 2 * kHalo, nzpad = G.nz + G.nb + 2 * kHalo, bx0 = G.nb + kHalo;
  const long long stride = ent_.size() > 1 ? ent_[1].f - ent_[0].f : (long long)nxpad * nzpad;
  ycorr_.resize((size_t)stride * ent_.size());
  float* y = ycorr_.get();
  // the models (1 / vel_t) of a single group lie model after model, the
  // entries shot after shot of each model
  const long long mstride = model_off_.size() > 1 ? model_off_[1] - model_off_[0] : 0;
  NetGeom g{nxpad, nzpad, stride, (int)ent_.size(), bx0, nxpad - bx0, kHalo, nzpad - bx0,
            rv_.get() + model_off_[0], mstride, (int)G.sx.size()};
  set_forward_hook([net, g, y](float* prev, float* cur, long long, int entries, hipStream_t s) {
    NetGeom h = g;
    h.entries = entries;
    net->correct(prev, h, y, s);
    net->correct(cur, h, y, s);
  });
}

void Batch::step_from(const float* prev_host, const float* cur_host, int it0, int nsteps, float* prev_out,
                      float* cur_out) {
  hipStream_t s = stream_.get();
  if (nsteps < 1 || it0 < 0) fail(kErrArg, "step_from needs it0 >= 0 and nsteps >= 1");
  long long o = 0;
  for (size_t c = 0; c < lay_.size(); c++) {
    const BatchEntry& e = ent_[run_of_[c]];
    const long long npts = (long long)e.nxpad * e.nzpad;
    SWFWI_HIP(hipMemcpyAsync(fa_.get() + e.f, prev_host + o, sizeof(float) * npts, hipMemcpyHostToDevice, s));
    SWFWI_HIP(hipMemcpyAsync(fb_.get() + e.f, cur_host + o, sizeof(float) * npts, hipMemcpyHostToDevice, s));
    o += npts;
  }
  BatchKernelArgs a = args(true);
  a.dcal = nullptr;
  const int n = (int)ent_.size();
  for (int k = 0; k < nsteps; k++) {
    const int it = it0 + k;
    int act = n;
    while (act > 0 && ent_[act - 1].nt <= it) act--;
    a.it_or_k = it;
    a.rot = k;
    a.nblocks = tile0_[act];
    launch_batch_fwd(a, kSnapNone, order_, tile_x_, s);
    if (hook_ && (it + 2) % kSnapEvery == 0) {  // the correction, as in forward runs
      float* cur = (k & 1) == 0 ? fa_.get() : fb_.get();
      float* prev = (k & 1) == 0 ? fb_.get() : fa_.get();
      hook_(prev, cur, ent_.size() > 1 ? ent_[1].f - ent_[0].f : 0, act, s);
    }
  }
  SWFWI_HIP(hipGetLastError());
  // the last launch (k = nsteps-1) wrote the newest field over the prev buffer
  const float* cur = ((nsteps - 1) & 1) == 0 ? fa_.get() : fb_.get();
  const float* prev = ((nsteps - 1) & 1) == 0 ? fb_.get() : fa_.get();
  o = 0;
  for (size_t c = 0; c < lay_.size(); c++) {
    const BatchEntry& e = ent_[run_of_[c]];
    const long long npts = (long long)e.nxpad * e.nzpad;
    if (prev_out)
      SWFWI_HIP(hipMemcpyAsync(prev_out + o, prev + e.f, sizeof(float) * npts, hipMemcpyDeviceToHost, s));
    if (cur_out) SWFWI_HIP(hipMemcpyAsync(cur_out + o, cur + e.f, sizeof(float) * npts, hipMemcpyDeviceToHost, s));
    o += npts;
  }
  SWFWI_HIP(hipStreamSynchronize(s));
}

static void timed(hipStream_t s, double& ms_out, const std::function<void()>& body) {
  hipEvent_t e0, e1;
  SWFWI_HIP(hipEventCreate(&e0));
  SWFWI_HIP(hipEventCreate(&e1));
  SWFWI_HIP(hipEventRecord(e0, s));
  body();
  SWFWI_HIP(hipEventRecord(e1, s));
  SWFWI_HIP(hipEventSynchronize(e1));
  float ms = 0;
  SWFWI_HIP(hipEventElapsedTime(&ms, e0, e1));
  ms_out = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

void Batch::forward(float* dcal_host) {
  hipStream_t s = stream_.get();
  if (tune_ && !tuned_ && !fixed_x_) tune();
  auto body = [&] {
    if (!graphs_) {
      run_loop(true);
      return;
    }
    if (!graph_[0]) {  // the whole loop (fills included, as kernels) in one graph
      hipGraph_t g = nullptr;
      SWFWI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      run_loop(true);
      SWFWI_HIP(hipStreamEndCapture(s, &g));
      SWFWI_HIP(hipGraphInstantiate(&graph_[0], g, nullptr, nullptr, 0));
      (void)hipGraphDestroy(g);
    }
    SWFWI_HIP(hipGraphLaunch(graph_[0], s));
  };
  timed(s, last_ms[0], body);
  SWFWI_HIP(hipGetLastError());
  if (dcal_host) {
    long long o = 0;
    for (const Layout& L : lay_) {
      const long long m = (long long)L.nt * L.ng;
      SWFWI_HIP(hipMemcpyAsync(dcal_host + o, dcal_.get() + L.dcal, sizeof(float) * m, hipMemcpyDeviceToHost, s));
      o += m;
    }
  }
  SWFWI_HIP(hipStreamSynchronize(s));
}

void Batch::adjoint_from_observed(const float* dobs_host, float* vsrc_host) {
  hipStream_t s = stream_.get();
  long long o = 0;
  for (const Layout& L : lay_) {  // observed data into vsrc_, then the residual in place
    const long long m = (long long)L.nt * L.ng;
    SWFWI_HIP(hipMemcpyAsync(vsrc_.get() + L.dcal, dobs_host + o, sizeof(float) * m, hipMemcpyHostToDevice, s));
    o += m;
  }
  // the residual dobs - dcal replaces the recording, then transVsrc into vsrc_
  launch_sub(vsrc_.get(), dcal_.get(), dcal_.get(), dcal_len_, s);
  // transVsrc per group: a group's entries are contiguous in constructor order
  size_t i = 0;
  while (i < lay_.size()) {
    size_t j = i;
    while (j < lay_.size() && lay_[j].group == lay_[i].group) j++;
    const Layout& L = lay_[i];
    const long long stride = round64((long long)L.nt * L.ng);
    if (stride != (long long)L.nt * L.ng) {
      for (size_t k = i; k < j; k++)
        op_trans_vsrc(dcal_.get() + lay_[k].dcal, vsrc_.get() + lay_[k].dcal, L.nt, L.ng, 1, s);
    } else {
      op_trans_vsrc(dcal_.get() + L.dcal, vsrc_.get() + L.dcal, L.nt, L.ng, (int)(j - i), s);
    }
    i = j;
  }
  if (vsrc_host) {
    o = 0;
    for (const Layout& L : lay_) {
      const long long m = (long long)L.nt * L.ng;
      SWFWI_HIP(hipMemcpyAsync(vsrc_host + o, vsrc_.get() + L.dcal, sizeof(float) * m, hipMemcpyDeviceToHost, s));
      o += m;
    }
  }
  SWFWI_HIP(hipStreamSynchronize(s));
}

void Batch::backward(const float* vsrc_host, float* grad_host, float* model_grad_host) {
  hipStream_t s = stream_.get();
  if (vsrc_host) {
    long long o = 0;
    for (const Layout& L : lay_) {
      const long long m = (long long)L.nt * L.ng;
      SWFWI_HIP(hipMemcpyAsync(vsrc_.get() + L.dcal, vsrc_host + o, sizeof(float) * m, hipMemcpyHostToDevice, s));
      o += m;
    }
  }
  auto body = [&] {
    if (!graphs_) {
      run_loop(false);
      return;
    }
    if (!graph_[1]) {
      hipGraph_t g = nullptr;
      SWFWI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      run_loop(false);
      SWFWI_HIP(hipStreamEndCapture(s, &g));
      SWFWI_HIP(hipGraphInstantiate(&graph_[1], g, nullptr, nullptr, 0));
      (void)hipGraphDestroy(g);
    }
    SWFWI_HIP(hipGraphLaunch(graph_[1], s));
  };
  timed(s, last_ms[1], body);
  SWFWI_HIP(hipGetLastError());
  // per-model sums over shots, in shot order (a model's shots are contiguous)
  size_t i = 0;
  int mi = 0;
  while (i < lay_.size()) {
    size_t j = i;
    while (j < lay_.size() && lay_[j].group == lay_[i].group && lay_[j].model == lay_[i].model) j++;
    const Layout& L = lay_[i];
    const long long npts = (long long)L.nxpad * L.nzpad;
    op_sum_entries(grad_.get() + L.grad, round64(npts), (int)(j - i), npts, mgrad_.get() + model_off_[mi], s);
    i = j;
    mi++;
  }
  if (grad_host) {
    long long o = 0;
    for (const Layout& L : lay_) {
      const long long npts = (long long)L.nxpad * L.nzpad;
      SWFWI_HIP(hipMemcpyAsync(grad_host + o, grad_.get() + L.grad, sizeof(float) * npts, hipMemcpyDeviceToHost, s));
      o += npts;
    }
  }
  if (model_grad_host) {
    long long o = 0;
    for (size_t m = 0; m < model_off_.size(); m++) {
      const BatchGroup& G = g_[model_group_[m]];
      const long long npts = (long long)(G.nx + 2 * G.nb + 2 * kHalo) * (G.nz + G.nb + 2 * kHalo);
      SWFWI_HIP(hipMemcpyAsync(model_grad_host + o, mgrad_.get() + model_off_[m], sizeof(float) * npts,
                               hipMemcpyDeviceToHost, s));
      o += npts;
    }
  }
  SWFWI_HIP(hipStreamSynchronize(s));
}

double Batch::forward_points() const {
  double p = 0;
  for (const BatchEntry& e : ent_) p += (double)e.nt * (e.nxpad - 2 * kHalo) * (e.nzpad - 2 * kHalo);
  return p;
}

double Batch::backward_points() const {
  double p = 0;
  for (const BatchEntry& e : ent_)
    p += (double)(e.nt - e.it_last) * (e.nxpad - 2 * kHalo) * (e.nzpad - 2 * kHalo);
  return p;
}

}  // namespace swfwi
