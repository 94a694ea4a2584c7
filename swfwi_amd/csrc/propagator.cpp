#include "propagator.h"

#include <algorithm>
#include <map>
#include <numeric>
#include <string>

namespace swfwi {

void Geometry::derive() {
  if (nx <= 0 || nz <= 0 || nt <= 0) fail(kErrArg, "grid and time dimensions must be positive");
  if (nb < 6) fail(kErrArg, "sponge width nb must be >= 6 (the 6-wide boundary band lives inside the sponge)");
  if (!(dt > 0) || !(dx > 0)) fail(kErrArg, "dt and dx must be positive");
  if ((long long)(nx + 2 * nb + 12) * (nz + nb + 12) >= (1LL << 30))
    fail(kErrRange, "padded grid too large: a field is indexed with 32-bit byte offsets");
  if (sx.size() != sz.size() || gx.size() != gz.size()) fail(kErrArg, "position arrays differ in length");
  if (sx.empty() || gx.empty()) fail(kErrArg, "need at least one source and one receiver");
  // Params::check (main-fm-damp.cpp:183-194): every position inside the model
  for (size_t i = 0; i < sx.size(); i++)
    if (sx[i] < 0 || sz[i] < 0 || sx[i] >= nx || sz[i] >= nz) fail(kErrArg, "sources exceeds the computing zone");
  for (size_t i = 0; i < gx.size(); i++)
    if (gx[i] < 0 || gz[i] < 0 || gx[i] >= nx || gz[i] >= nz) fail(kErrArg, "geophones exceeds the computing zone");
  nxpad = nx + 2 * nb + 2 * kHalo;
  nzpad = nz + nb + 2 * kHalo;
  ld = nzpad;
  bz0 = kHalo;
  bx0 = bxn = bzn = nb + kHalo;
  const int w = 6;
  const int nxb = nxpad - (bx0 - w + bxn - w);
  const int nzb = nzpad - bz0 - bzn;
  bsize = w * (nxb + 2 * nzb);
  ntz = (nzpad - 2 * kHalo + kTZ - 1) / kTZ;
  ntx = tile_count_x(nxpad, 16);
}

void build_damping(int nx, int nz, int nb, DampTables& t, hipStream_t s) {
  // fd4t10s-damp-zjh.c:64-80 with bx = bz = nb; the bottom branch wins.
  const float max_delta = 0.05f;
  std::vector<float> dxv(nx), dzv(nz, 0.f);
  for (int ix = 0; ix < nx; ix++) {
    float dist = 0;
    if (ix < nb) dist = (float)(nb - ix) / nb;
    if (ix >= nx - nb) dist = (float)(ix - (nx - nb) + 1) / nb;
    dxv[ix] = max_delta * dist * dist;
  }
  for (int iz = 0; iz < nz; iz++) {
    if (iz >= nz - nb) {
      float dist = (float)(iz - (nz - nb) + 1) / nb;
      dzv[iz] = max_delta * dist * dist;
    }
  }
  t.dx.upload(dxv, s);
  t.dz.upload(dzv, s);
  t.zdamp = nz - nb;
  SWFWI_HIP(hipStreamSynchronize(s));  // the host vectors end here
}

void build_points(const std::vector<long long>& idx, int ld, int ntz, int ntx, int tile_x, int halo, DevPoints& out,
                  hipStream_t s) {
  // tile t = tx*ntz + tz covers outputs [kHalo + tx*tile_x, +tile_x) x
  // [kHalo + tz*kTZ, +kTZ); a point is listed under every tile whose cover
  // grown by `halo` cells holds it, in ascending point id
  const int ntiles = ntz * ntx;
  const int n = (int)idx.size();
  std::vector<std::vector<int>> lists(ntiles);
  for (int p = 0; p < n; p++) {
    const int ix = (int)(idx[p] / ld), iz = (int)(idx[p] % ld);
    const int tz = (iz - kHalo) / kTZ, tx = (ix - kHalo) / tile_x;
    if (iz < kHalo || ix < kHalo || tz >= ntz || tx >= ntx) fail(kErrArg, "point outside the stencil update region");
    const int z_lo = std::max(0, (iz - kHalo - halo) / kTZ - 1), z_hi = std::min(ntz - 1, tz + 1 + halo / kTZ);
    const int x_lo = std::max(0, (ix - kHalo - halo) / tile_x - 1), x_hi = std::min(ntx - 1, tx + 1 + halo / tile_x);
    for (int x = x_lo; x <= x_hi; x++)
      for (int z = z_lo; z <= z_hi; z++) {
        const int gx0 = kHalo + x * tile_x, gz0 = kHalo + z * kTZ;
        if (ix >= gx0 - halo && ix < gx0 + tile_x + halo && iz >= gz0 - halo && iz < gz0 + kTZ + halo)
          lists[(size_t)x * ntz + z].push_back(p);
      }
  }
  std::vector<int> ptr(ntiles + 1, 0), items;
  for (int t = 0; t < ntiles; t++) {
    ptr[t + 1] = ptr[t] + (int)lists[t].size();
    items.insert(items.end(), lists[t].begin(), lists[t].end());
  }
  std::vector<int> idx32(n), prev_same(n, -1), next_same(n, -1);
  for (int p = 0; p < n; p++) idx32[p] = (int)idx[p];
  {
    std::map<int, int> last;  // grid index -> last point id seen there
    for (int p = 0; p < n; p++) {
      auto it = last.find(idx32[p]);
      if (it != last.end()) {
        prev_same[p] = it->second;
        next_same[it->second] = p;
      }
      last[idx32[p]] = p;
    }
  }
  out.prev_same.upload(prev_same, s);
  out.next_same.upload(next_same, s);
  out.tile_ptr.upload(ptr, s);
  out.items.upload(items, s);
  out.idx.upload(idx32, s);
  out.n = n;
  out.max_per_tile = 0;
  for (int t = 0; t < ntiles; t++) out.max_per_tile = std::max(out.max_per_tile, ptr[t + 1] - ptr[t]);
  std::vector<int> count(ntiles);
  std::vector<int4> slots;
  if (halo == 0 && out.max_per_tile <= kTilePts) {
    // (grid index, id, position inside the tile, shares its grid index with
    // another point of the tile)
    slots.assign((size_t)ntiles * kTilePts, make_int4(-1, 0, 0, 0));
    for (int t = 0; t < ntiles; t++) {
      count[t] = ptr[t + 1] - ptr[t];
      const int x0 = kHalo + (t / ntz) * tile_x, z0 = kHalo + (t % ntz) * kTZ;
      for (int j = ptr[t]; j < ptr[t + 1]; j++) {
        const int pid = items[j], q = idx32[pid];
        int shared = 0;
        for (int k = ptr[t]; k < ptr[t + 1]; k++)
          if (k != j && idx32[items[k]] == q) shared = 1;
        const int loc = (q / ld - x0) * kTZ + (q % ld - z0);
        slots[(size_t)t * kTilePts + (j - ptr[t])] = make_int4(q, pid, loc, shared);
      }
    }
    out.count.upload(count, s);
    out.slots.upload(slots, s);
  }
  SWFWI_HIP(hipStreamSynchronize(s));  // the host vectors end here
}

void Fields::ensure(int e, long long npts, bool receiver) {
  E = e;
  const size_t n = (size_t)e * (size_t)npts;
  a.resize(n);
  b.resize(n);
  if (receiver) {
    c.resize(n);
    d.resize(n);
  }
}

Propagator::Propagator(const Geometry& g, int precision) : g_(g), prec_(precision) {
  g_.derive();
  if (precision != kExact && precision != kFp32) fail(kErrArg, "unknown precision mode");
  SWFWI_HIP(hipGetDevice(&dev_));
  for (hipEvent_t& e : lane_events_) SWFWI_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  hipStream_t s = stream_.get();
  build_damping(g_.nxpad, g_.nzpad, g_.bx0, damp_, s);
  std::vector<long long> si(g_.ns()), gi(g_.ng());
  std::vector<int> sxp(g_.ns()), szp(g_.ns()), gxp(g_.ng()), gzp(g_.ng());
  for (int i = 0; i < g_.ns(); i++) {
    sxp[i] = g_.sx[i] + g_.bx0;
    szp[i] = g_.sz[i] + g_.bz0;
    si[i] = (long long)sxp[i] * g_.ld + szp[i];
  }
  for (int i = 0; i < g_.ng(); i++) {
    gxp[i] = g_.gx[i] + g_.bx0;
    gzp[i] = g_.gz[i] + g_.bz0;
    gi[i] = (long long)gxp[i] * g_.ld + gzp[i];
  }
  for (int w = 0; w < kNumWidths; w++) {
    const int ntx = tile_count_x(g_.nxpad, kTileWidths[w]);
    build_points(si, g_.ld, g_.ntz, ntx, kTileWidths[w], 0, src_[w], s);
    build_points(gi, g_.ld, g_.ntz, ntx, kTileWidths[w], 0, rec_[w], s);
  }
  sxp_.upload(sxp, s);
  szp_.upload(szp, s);
  gxp_.upload(gxp, s);
  gzp_.upload(gzp, s);
  stream_.sync();
}

int Propagator::block_chunk(int kind, int E, long long rvstride, int rvdiv) const {
  // Measured on the 301x921 grid: below 4 entries the order does not matter
  // (latency-bound).  A forward step whose entries share a model is fastest
  // entries first (the model, a quarter of its traffic, read once: 60
  // entries 65 against 70 us chunked); with one model per entry it is
  // fastest in chunks of 4 tile columns (ensemble epoch of 100 members 3.05
  // -> 2.82 s).  The fused backward iteration (four per-entry fields) runs
  // chunked (100 entries 286 -> 262 us), the split one entries first (30
  // entries 73 against 79 us).
  if (E < 4) return 0;
  const bool chunked = kind == 0 ? (rvstride != 0 && rvdiv == 1) : !bwd_split(E);
  return chunked ? 4 * g_.ntz : 0;
}

Grid Propagator::grid(int w) const {
  Grid gr;
  gr.nx = g_.nxpad;
  gr.nz = g_.nzpad;
  gr.ld = g_.ld;
  gr.ntz = g_.ntz;
  gr.tile_x = kTileWidths[w];
  gr.ntx = tile_count_x(g_.nxpad, gr.tile_x);
  gr.nb = g_.bx0;
  gr.bx0 = g_.bx0;
  gr.bxn = g_.bxn;
  gr.bz0 = g_.bz0;
  gr.bzn = g_.bzn;
  gr.delta_x = damp_.dx.get();
  gr.delta_z = damp_.dz.get();
  gr.zdamp = damp_.zdamp;
  return gr;
}

void Propagator::set_bwd_split(int mode) {
  if (mode < -1 || mode > 1) fail(kErrArg, "backward split mode must be -1 (by batch size), 0 or 1");
  bwd_split_ = mode;
}

bool Propagator::bwd_split(int E) const {
  if (bwd_split_ >= 0) return bwd_split_ == 1;
  // measured on the 301x921 grid with the imaging condition: split (64x16
  // tiles, entries fastest) up to 16 entries; fused (64x16, in chunks of tile
  // columns) from 30 entries on, where the split form's second read of the
  // receiver field costs more than the fused form's occupancy
  const long long tiles16 = (long long)g_.ntz * tile_count_x(g_.nxpad, 16) * E;
  return tiles16 < kBwdSplitBelow;
}

void Propagator::set_tile_x(int tx) {
  if (tx != 0 && tx != 8 && tx != 16 && tx != 32) fail(kErrArg, "tile width must be 0 (by rule), 8, 16 or 32");
  tile_override_ = tx;
}

long long Propagator::tile_key(int kind, int E) const {
  return ((long long)kind << 40) + ((long long)(kind == 1 && bwd_split(E)) << 32) + E;
}

int Propagator::tile_choice_cached(int kind, int E) const {
  if (tile_override_) return tile_override_;
  const long long key = tile_key(kind, E);
  for (const auto& kv : tile_cache_)
    if (kv.first == key) return kTileWidths[kv.second];
  return 0;
}

float Propagator::time_steps(int kind, int E, const float* rv, long long rvstride, int rvdiv, Fields& w, int width,
                             int reps, bool clear) {
  // `reps` back-to-back launches of one step kernel on scratch wavefields
  // (values discarded), timed with one event pair; microseconds per launch.
  // The backward iteration includes the imaging condition (into a scratch
  // gradient), which decides between its fused and split forms.
  hipStream_t s = stream();
  const long long np = g_.npts();
  w.ensure(E, np, true);
  DBuf<float> own;
  float* gscratch = nullptr;
  if (kind == 1) {
    const size_t n = (size_t)E * (size_t)np;
    if (tune_grad_.size() >= n) {
      gscratch = tune_grad_.get();
    } else {
      own.resize(n);
      gscratch = own.get();
    }
    launch_fill(gscratch, (long long)n, 0.f, s);
  }
  if (clear) {
    launch_fill(w.a.get(), (long long)E * np, 0.f, s);
    launch_fill(w.b.get(), (long long)E * np, 0.f, s);
    launch_fill(w.c.get(), (long long)E * np, 0.f, s);
    launch_fill(w.d.get(), (long long)E * np, 0.f, s);
  }
  hipEvent_t e0, e1;
  SWFWI_HIP(hipEventCreate(&e0));
  SWFWI_HIP(hipEventCreate(&e1));
  SWFWI_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < reps; r++) {
    if (kind == 0) {
      FwdStepArgs a;
      a.g = grid(width);
      a.fstride = np;
      a.cur = (r & 1) ? w.a.get() : w.b.get();
      a.prev = (r & 1) ? w.b.get() : w.a.get();
      a.rv = rv;
      a.rvstride = rvstride;
      a.rvdiv = rvdiv;
      launch_fwd_step(a, E, true, prec_, s, block_chunk(0, E, rvstride, rvdiv));
    } else {
      BwdStepArgs a;
      a.g = grid(width);
      a.fstride = np;
      a.s_cur = (r & 1) ? w.a.get() : w.b.get();
      a.s_prev = (r & 1) ? w.b.get() : w.a.get();
      a.g_cur = (r & 1) ? w.c.get() : w.d.get();
      a.g_prev = (r & 1) ? w.d.get() : w.c.get();
      a.rv = rv;
      a.rvstride = rvstride;
      a.rvdiv = rvdiv;
      a.grad = gscratch;
      a.gstride = np;
      a.weight = 1.f;
      a.do_xcorr = 1;
      a.split = bwd_split(E);
      launch_bwd_step(a, E, prec_, s, block_chunk(1, E, rvstride, rvdiv));
    }
  }
  SWFWI_HIP(hipEventRecord(e1, s));
  SWFWI_HIP(hipEventSynchronize(e1));
  float ms = 0;
  SWFWI_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  SWFWI_HIP(hipGetLastError());
  return 1e3f * ms / (reps > 0 ? reps : 1);
}

int Propagator::fixed_width(int kind, int E) const {
  if (tile_override_) {
    for (int i = 0; i < kNumWidths; i++)
      if (kTileWidths[i] == tile_override_) return i;
  }
  // measured on the 301x921 grid (372 tiles of 64x16 per entry): the narrow
  // tile wins while the batch does not fill the chip (more, shorter blocks),
  // the wide one once it is bandwidth-bound (less halo traffic per output).
  // Forward: 64x8 at one entry, 64x16 up to 6, 64x32 beyond; backward: 64x8
  // at one entry, 64x16 beyond (split up to 16 entries, fused beyond)
  const long long tiles16 = (long long)g_.ntz * tile_count_x(g_.nxpad, 16) * E;
  if (kind == 1) return tiles16 < 600 ? 0 : 1;
  return tiles16 < 600 ? 0 : (tiles16 < 2500 ? 1 : 2);
}

int Propagator::tile_choice(int kind, int E, const float* rv, long long rvstride, int rvdiv, Fields& w,
                            bool may_time) {
  if (tile_override_) return fixed_width(kind, E);
  const long long key = tile_key(kind, E);
  for (const auto& kv : tile_cache_)
    if (kv.first == key) return kv.second;
  const int rule = fixed_width(kind, E);  // the measured rule is the incumbent
  int best = rule;
  if (tune_ && may_time) {
    // timed trial on scratch wavefields (the caller's workspace is free at
    // this point: forward() tunes before it zeroes and runs); three
    // interleaved rounds, the first warming clocks and caches; a width
    // replaces the rule's only when it is faster by kTuneMargin in both of
    // the last two, so that noise between near-equal widths keeps the rule
    float t[2][kNumWidths] = {};
    if (kind == 1) tune_grad_.resize((size_t)E * (size_t)g_.npts());
    for (int round = 0; round < 3; round++)
      for (int i = 0; i < kNumWidths; i++) {
        const float us = time_steps(kind, E, rv, rvstride, rvdiv, w, i, 8, round == 0 && i == 0);
        if (round > 0) t[round - 1][i] = us;
      }
    tune_grad_ = DBuf<float>();
    for (int i = 0; i < kNumWidths; i++) {
      const bool faster = t[0][i] < (1.f - kTuneMargin) * t[0][rule] && t[1][i] < (1.f - kTuneMargin) * t[1][rule];
      if (faster && std::max(t[0][i], t[1][i]) < std::max(t[0][best], t[1][best])) best = i;
    }
  }
  tile_cache_.push_back({key, best});
  return best;
}

int Propagator::xcorr_range(float dt, int nt, std::vector<float>& weights) {
  // fwiframework.cpp:164-175: float dt*it against double thresholds; the loop
  // breaks at the first it without an imaging condition.
  weights.assign(nt, 0.f);
  int it_min = nt;
  for (int it = nt - 1; it >= 0; it--) {
    if (dt * it > 0.4) {
      weights[it] = 1.0;
    } else if (dt * it > 0.3) {
      weights[it] = (dt * it - 0.3) / 0.1;
    } else {
      break;
    }
    it_min = it;
  }
  return it_min;
}

std::vector<Propagator::FwdLaunch> Propagator::plan_forward(Fields& fw) const {
  // Buffer rotation of a forward run: b holds p(0) (the injected zero field),
  // a holds p(-1) = 0; each step writes next over prev.
  const int nt = g_.nt;
  std::vector<FwdLaunch> plan;
  float* cur = fw.b.get();
  float* prv = fw.a.get();
  for (int it = 0; it < nt; it++) {
    plan.push_back({it, cur, prv});
    std::swap(cur, prv);
  }
  // after the last step: s_cur = p(nt-1), s_nxt = p(nt)
  fw.s_cur = prv;
  fw.s_nxt = cur;
  fw.spare0 = fw.c.get();
  fw.spare1 = fw.d.get();
  return plan;
}

void Propagator::forward_body(int E, const float* rv, long long rvstride, const Sources& src, float* dcal,
                              float* band, float* fa, float* fb, int rvdiv, int w, hipStream_t s) {
  const long long np = g_.npts();
  const int nt = g_.nt, ng = g_.ng();
  launch_fill(fa, (long long)E * np, 0.f, s);
  launch_fill(fb, (long long)E * np, 0.f, s);

  Inject inj;
  inj.pts = src_[w].view();
  inj.amp = src.amp;
  inj.ld = src.ld;
  inj.estride = src.estride;
  inj.sign = +1;
  inj.lo = src.lo;
  inj.hi = src.hi;

  // step 0 injects into the initially zero cur (addSource before stepForward)
  inj.t = 0;
  launch_point_inject(fb, np, inj, g_.ns(), E, s);

  FwdStepArgs a;
  a.g = grid(w);
  a.fstride = np;
  a.rv = rv;
  a.rvstride = rvstride;
  a.rvdiv = rvdiv;
  a.inj = inj;
  a.rec.pts = rec_[w].view();
  a.rec.estride = (long long)nt * ng;
  a.band_estride = (long long)nt * g_.bsize;
  float* cur = fb;
  float* prv = fa;
  for (int it = 0; it < nt; it++, std::swap(cur, prv)) {
    a.cur = cur;
    a.prev = prv;
    a.inj.t = it + 1 < nt ? it + 1 : -1;
    a.rec.out = dcal ? dcal + (long long)it * ng : nullptr;
    a.band = band ? band + (long long)it * g_.bsize : nullptr;
    launch_fwd_step(a, E, true, prec_, s, block_chunk(0, E, rvstride, rvdiv));
  }
  SWFWI_HIP(hipGetLastError());
}

void Propagator::forward(int E, const float* rv, long long rvstride, const Sources& src, float* dcal, float* band,
                         Fields& w, int rvdiv) {
  const long long np = g_.npts();
  w.ensure(E, np, true);
  // tile widths for this batch, chosen now while the workspace is free
  // (backward() runs on the state this forward leaves behind)
  const int tw = tile_choice(0, E, rv, rvstride, rvdiv, w, true);
  tile_choice(1, E, rv, rvstride, rvdiv, w, true);
  const uintptr_t key[] = {1,
                           (uintptr_t)tw,
                           (uintptr_t)E,
                           (uintptr_t)rv,
                           (uintptr_t)rvstride,
                           (uintptr_t)rvdiv,
                           (uintptr_t)src.amp,
                           (uintptr_t)src.ld,
                           (uintptr_t)src.estride,
                           (uintptr_t)src.lo,
                           (uintptr_t)src.hi,
                           (uintptr_t)dcal,
                           (uintptr_t)band,
                           (uintptr_t)w.a.get(),
                           (uintptr_t)w.b.get(),
                           (uintptr_t)w.c.get(),
                           (uintptr_t)w.d.get(),
                           (uintptr_t)prec_};
  const bool lanes = E >= 2 && E <= kLanes && band == nullptr;
  if (!lanes) {
    hipStream_t s = stream();
    run_maybe_graph(
        std::vector<uintptr_t>(std::begin(key), std::end(key)), E, 0, g_.nt,
        [&] { forward_body(E, rv, rvstride, src, dcal, band, w.a.get(), w.b.get(), rvdiv, tw, s); }, s);
  } else {
    // entry e as a one-entry run on lane e, over its own slice of the batch's
    // buffers (the same rotation, so the batch's final state is unchanged);
    // the lanes start after the work queued so far and the main stream
    // continues after all of them
    const int tw1 = tile_choice(0, 1, rv, rvstride, rvdiv, w, false);
    hipStream_t main = stream();
    SWFWI_HIP(hipEventRecord(lane_events_[0], main));
    for (int e = 0; e < E; e++) {
      hipStream_t ls = e == 0 ? main : lane_streams_[e - 1].get();
      if (e > 0) SWFWI_HIP(hipStreamWaitEvent(ls, lane_events_[0], 0));
      Sources se = src;
      se.amp = src.amp + (long long)e * src.estride;
      se.lo = src.lo ? src.lo + e : nullptr;
      se.hi = src.hi ? src.hi + e : nullptr;
      const float* rve = rv + (long long)(e / rvdiv) * rvstride;
      float* dce = dcal ? dcal + (long long)e * g_.nt * g_.ng() : nullptr;
      float* fa = w.a.get() + (long long)e * np;
      float* fb = w.b.get() + (long long)e * np;
      const uintptr_t lkey[] = {3, (uintptr_t)tw1, (uintptr_t)e, (uintptr_t)ls, (uintptr_t)rve,
                                (uintptr_t)se.amp, (uintptr_t)src.ld, (uintptr_t)se.lo, (uintptr_t)se.hi,
                                (uintptr_t)dce, (uintptr_t)fa, (uintptr_t)fb, (uintptr_t)prec_};
      run_maybe_graph(std::vector<uintptr_t>(std::begin(lkey), std::end(lkey)), 1, 0, g_.nt,
                      [&] { forward_body(1, rve, 0, se, dce, nullptr, fa, fb, 1, tw1, ls); }, ls);
    }
    for (int e = 1; e < E; e++) {
      SWFWI_HIP(hipEventRecord(lane_events_[e], lane_streams_[e - 1].get()));
      SWFWI_HIP(hipStreamWaitEvent(main, lane_events_[e], 0));
    }
  }
  const std::vector<FwdLaunch> plan = plan_forward(w);  // sets the final state
  flush_timing_if_full();
  fwd_launches += (long long)plan.size() * (lanes ? E : 1);
  fwd_point_updates += (double)E * (g_.nxpad - 2 * kHalo) * (g_.nzpad - 2 * kHalo) * g_.nt;
}

void Propagator::forward_external(const float* rv, const Sources& src, float* dcal, float* fa, float* fb) {
  forward_body(1, rv, 0, src, dcal, nullptr, fa, fb, 1, fixed_width(0, 1), stream());
  fwd_launches += g_.nt;
  fwd_point_updates += (double)(g_.nxpad - 2 * kHalo) * (g_.nzpad - 2 * kHalo) * g_.nt;
}

void Propagator::set_timing(bool on) {
  collect_timing();
  timing_ = on;
  kernel_ms[0] = kernel_ms[1] = 0;
  kernel_n[0] = kernel_n[1] = 0;
}

void Propagator::collect_timing() {
  // timed loops still in flight: wait for each loop's own stop event (a loop
  // may run on a lane stream that has not joined the main stream yet) and
  // add their durations
  if (pending_.empty()) return;
  for (const TimedLoop& t : pending_) {
    float ms = 0;
    SWFWI_HIP(hipEventSynchronize(t.stop));
    SWFWI_HIP(hipEventElapsedTime(&ms, t.start, t.stop));
    kernel_ms[t.kind] += ms;
    kernel_n[t.kind] += t.launches;
    event_pool_.push_back(t.start);
    event_pool_.push_back(t.stop);
  }
  pending_.clear();
}

hipEvent_t Propagator::pooled_event() {
  hipEvent_t e;
  if (!event_pool_.empty()) {
    e = event_pool_.back();
    event_pool_.pop_back();
  } else {
    SWFWI_HIP(hipEventCreate(&e));
  }
  return e;
}

void Propagator::run_maybe_graph(const std::vector<uintptr_t>& key, int E, int kind, long long launches,
                                 const std::function<void()>& body, hipStream_t s) {
  // timing: one event pair around the whole time loop (outside any graph),
  // so the measured loop is the production one, launch gaps included
  hipEvent_t t0 = nullptr;
  if (timing_) {
    t0 = pooled_event();
    SWFWI_HIP(hipEventRecord(t0, s));
  }
  // Large batches are not launch-bound: a graph would only add its capture.
  if (graphs_ == 0 || (graphs_ == 1 && E > kGraphMaxEntries)) {
    body();
  } else {
    GraphEntry* hit = nullptr;
    for (auto& ge : graph_cache_)
      if (ge.key == key) hit = &ge;
    if (!hit) {
      hipGraph_t graph = nullptr;
      SWFWI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        body();
      } catch (...) {
        (void)hipStreamEndCapture(s, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      SWFWI_HIP(hipStreamEndCapture(s, &graph));
      hipGraphExec_t exec = nullptr;
      SWFWI_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      (void)hipGraphDestroy(graph);
      if (graph_cache_.size() >= 16) {
        (void)hipGraphExecDestroy(graph_cache_.front().exec);
        graph_cache_.erase(graph_cache_.begin());
      }
      graph_cache_.push_back(GraphEntry{key, exec});
      hit = &graph_cache_.back();
    }
    SWFWI_HIP(hipGraphLaunch(hit->exec, s));
  }
  if (timing_) {
    hipEvent_t t1 = pooled_event();
    SWFWI_HIP(hipEventRecord(t1, s));
    pending_.push_back(TimedLoop{t0, t1, kind, launches});
  }
}

void Propagator::flush_timing_if_full() {
  // called once a forward()/backward() has joined its lanes, so every pending
  // loop is ordered before work already queued on the main stream
  if (pending_.size() >= (size_t)timing_flush_) collect_timing();
}

Propagator::~Propagator() {
  for (auto& ge : graph_cache_) (void)hipGraphExecDestroy(ge.exec);
  for (const TimedLoop& t : pending_) {
    (void)hipEventDestroy(t.start);
    (void)hipEventDestroy(t.stop);
  }
  for (hipEvent_t e : event_pool_) (void)hipEventDestroy(e);
  for (hipEvent_t e : lane_events_)
    if (e) (void)hipEventDestroy(e);
}

int Propagator::backward(int E, const float* rv, long long rvstride, const Sources& src, const float* vsrc,
                         long long vstride, const float* band, float* grad, long long gstride, Fields& w,
                         int rvdiv) {
  const long long np = g_.npts();
  if (!w.s_cur) fail(kErrState, "backward() needs the state of a forward() run");
  w.ensure(E, np, true);
  std::vector<float> wts;
  const int it_min = xcorr_range(g_.dt, g_.nt, wts);
  if (it_min >= g_.nt) return 0;
  const int tw = tile_choice(1, E, rv, rvstride, rvdiv, w, false);
  const uintptr_t key[] = {2,
                           (uintptr_t)tw,
                           (uintptr_t)E,
                           (uintptr_t)rv,
                           (uintptr_t)rvstride,
                           (uintptr_t)rvdiv,
                           (uintptr_t)src.amp,
                           (uintptr_t)src.ld,
                           (uintptr_t)src.estride,
                           (uintptr_t)src.lo,
                           (uintptr_t)src.hi,
                           (uintptr_t)vsrc,
                           (uintptr_t)vstride,
                           (uintptr_t)band,
                           (uintptr_t)grad,
                           (uintptr_t)gstride,
                           (uintptr_t)w.s_cur,
                           (uintptr_t)w.s_nxt,
                           (uintptr_t)w.spare0,
                           (uintptr_t)w.spare1,
                           (uintptr_t)prec_,
                           (uintptr_t)bwd_split(E)};
  run_maybe_graph(std::vector<uintptr_t>(std::begin(key), std::end(key)), E, 1, g_.nt - it_min, [&] {
    backward_body(E, rv, rvstride, src, vsrc, vstride, band, grad, gstride, w, rvdiv, it_min, wts, tw);
  }, stream());
  flush_timing_if_full();
  const int iters = g_.nt - it_min;
  bwd_launches += iters;
  bwd_point_updates += 2.0 * E * (g_.nxpad - 2 * kHalo) * (g_.nzpad - 2 * kHalo) * iters;
  return iters;
}

void Propagator::backward_body(int E, const float* rv, long long rvstride, const Sources& src, const float* vsrc,
                               long long vstride, const float* band, float* grad, long long gstride, Fields& fw,
                               int rvdiv, int it_min, const std::vector<float>& wts, int w) {
  hipStream_t s = stream();
  const long long np = g_.npts();
  const int nt = g_.nt, ng = g_.ng();
  float* s_cur = fw.s_cur;
  float* s_prv = fw.s_nxt;
  float* g_cur = fw.spare0;
  float* g_prv = fw.spare1;
  launch_fill(g_cur, (long long)E * np, 0.f, s);
  launch_fill(g_prv, (long long)E * np, 0.f, s);

  Inject ginj;
  ginj.pts = rec_[w].view();
  ginj.amp = vsrc;
  ginj.ld = ng;
  ginj.estride = vstride;
  ginj.sign = +1;
  ginj.t = nt - 1;
  launch_point_inject(g_cur, np, ginj, ng, E, s);

  BwdStepArgs a;
  a.g = grid(w);
  a.fstride = np;
  a.rv = rv;
  a.rvstride = rvstride;
  a.rvdiv = rvdiv;
  a.s_sub.pts = src_[w].view();
  a.s_sub.amp = src.amp;
  a.s_sub.ld = src.ld;
  a.s_sub.estride = src.estride;
  a.s_sub.sign = -1;
  a.s_sub.lo = src.lo;
  a.s_sub.hi = src.hi;
  a.g_inj = ginj;
  a.band_estride = (long long)nt * g_.bsize;
  a.grad = grad;
  a.gstride = gstride;
  a.do_xcorr = 1;
  a.split = bwd_split(E);
  for (int it = nt - 1; it >= it_min; it--) {
    a.s_cur = s_cur;
    a.s_prev = s_prv;
    a.g_cur = g_cur;
    a.g_prev = g_prv;
    a.s_sub.t = it;
    const bool more = it - 1 >= it_min;
    a.g_inj.t = more ? it - 1 : -1;
    a.band = more ? band + (long long)(it - 1) * g_.bsize : nullptr;
    a.band_next = it - 2 >= it_min ? band + (long long)(it - 2) * g_.bsize : nullptr;
    a.weight = wts[it];
    launch_bwd_step(a, E, prec_, s, block_chunk(1, E, rvstride, rvdiv));
    std::swap(s_cur, s_prv);
    std::swap(g_cur, g_prv);
  }
  SWFWI_HIP(hipGetLastError());
}

}  // namespace swfwi
