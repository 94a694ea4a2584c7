// extern "C" surface of libswfwi_hip.so (declared in include/swfwi_hip.h).
#include "../../include/swfwi_hip.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "batch.h"
#include "hostmath.h"
#include "learned.h"
#include "noalloc.h"
#include "solver.h"

using namespace swfwi;

struct swfwi_solver_s {
  std::unique_ptr<Solver> s;
};
struct swfwi_inversion_s {
  std::unique_ptr<Inversion> inv;
};
struct swfwi_batch_s {
  std::unique_ptr<Batch> b;
};
struct swfwi_net_s {
  std::unique_ptr<ConvNet> n;
  std::unique_ptr<Stream> s;
};

namespace {

thread_local std::string g_last_error;

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return SWFWI_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return SWFWI_ERR_RANGE;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return SWFWI_ERR_ARG;
  }
}

#define NEED(p)                                             \
  do {                                                      \
    if (!(p)) fail(kErrArg, std::string("null argument: ") + #p); \
  } while (0)

// Device state of the per-step drop-in kernels: one set of buffers grown to
// the largest grid seen, guarded by a mutex (the reference's static u2 makes
// these calls non-reentrant; here they are merely serialised).
struct StepContext {
  std::mutex mu;
  std::unique_ptr<Stream> stream;
  DBuf<float> prev, cur, rv, u2;
  DampTables damp;
  int nx = -1, nz = -1, nb = -1;
  DevPoints none;
};
StepContext& step_ctx() {
  static StepContext c;
  return c;
}

int run_step(float* prev, const float* cur, const float* vel, float* u2, int nx, int nz, int nb, bool damped,
             int precision) {
  return guarded([&] {
    NEED(prev);
    NEED(cur);
    NEED(vel);
    if (nx < 2 * kHalo + 1 || nz < 2 * kHalo + 1) fail(kErrArg, "grid smaller than the stencil");
    if ((long long)nx * nz >= (1LL << 30)) fail(kErrRange, "grid too large: a field is indexed with 32-bit byte offsets");
    if (precision != kExact && precision != kFp32) fail(kErrArg, "unknown precision mode");
    StepContext& c = step_ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    if (!c.stream) c.stream.reset(new Stream());
    hipStream_t s = c.stream->get();
    const long long n = (long long)nx * nz;
    if (damped && (nx != c.nx || nz != c.nz || nb != c.nb)) {
      build_damping(nx, nz, nb, c.damp, s);
      c.nx = nx;
      c.nz = nz;
      c.nb = nb;
    }
    c.prev.upload(prev, (size_t)n, s);
    c.cur.upload(cur, (size_t)n, s);
    c.rv.upload(vel, (size_t)n, s);
    launch_reciprocal(c.rv.get(), c.rv.get(), n, s);
    c.u2.resize((size_t)n);
    Grid g;
    g.nx = nx;
    g.nz = nz;
    g.ld = nz;
    g.ntz = (nz - 2 * kHalo + kTZ - 1) / kTZ;
    g.ntx = tile_count_x(nx, 16);
    g.tile_x = 16;
    g.nb = nb;
    if (damped) {
      g.delta_x = c.damp.dx.get();
      g.delta_z = c.damp.dz.get();
      g.zdamp = c.damp.zdamp;
    }
    FwdStepArgs a;
    a.g = g;
    a.cur = c.cur.get();
    a.prev = c.prev.get();
    a.rv = c.rv.get();
    a.u2out = u2 ? c.u2.get() : nullptr;
    launch_fwd_step(a, 1, damped, precision, s);
    SWFWI_HIP(hipGetLastError());
    c.prev.download(prev, (size_t)n, s);
    std::vector<float> hu2;
    if (u2) {
      hu2.resize((size_t)n);
      c.u2.download(hu2.data(), (size_t)n, s);
    }
    SWFWI_HIP(hipStreamSynchronize(s));
    if (u2) {  // the reference writes u2 over [5, n-5) only
      for (int ix = 5; ix < nx - 5; ix++)
        for (int iz = 5; iz < nz - 5; iz++) u2[(size_t)ix * nz + iz] = hu2[(size_t)ix * nz + iz];
    }
  });
}

}  // namespace

extern "C" {

const char* swfwi_last_error(void) { return g_last_error.c_str(); }
const char* swfwi_version(void) { return "swfwi-mi355x 0.1 (gfx950)"; }

int swfwi_device_count(int* count) {
  return guarded([&] {
    NEED(count);
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
  });
}

int swfwi_copy_bandwidth(long long n, int iters, double* gbs) {
  return guarded([&] {
    NEED(gbs);
    if (n < 1024 || iters < 1) fail(kErrArg, "copy probe needs n >= 1024 and iters >= 1");
    Stream st;
    *gbs = op_copy_bandwidth(n, iters, st.get());
  });
}

int swfwi_launch_latency(int blocks, int reps, int graph, double* us) {
  return guarded([&] {
    NEED(us);
    if (blocks < 1 || reps < 1) fail(kErrArg, "blocks and reps must be >= 1");
    Stream st;
    *us = op_launch_latency(blocks, reps, graph != 0, st.get());
  });
}

void fd4t10s_damp_zjh_2d_vtrans(float* prev_wave, const float* curr_wave, const float* vel, float* u2, int nx, int nz,
                                int nb) {
  if (run_step(prev_wave, curr_wave, vel, u2, nx, nz, nb, true, kExact) != SWFWI_OK) {
    std::fprintf(stderr, "fd4t10s_damp_zjh_2d_vtrans: %s\n", g_last_error.c_str());
    std::exit(1);
  }
}

void fd4t10s_zjh_2d_vtrans(float* prev_wave, const float* curr_wave, const float* vel, float* u2, int nx, int nz) {
  if (run_step(prev_wave, curr_wave, vel, u2, nx, nz, 0, false, kExact) != SWFWI_OK) {
    std::fprintf(stderr, "fd4t10s_zjh_2d_vtrans: %s\n", g_last_error.c_str());
    std::exit(1);
  }
}

int swfwi_step_damped(float* prev_wave, const float* curr_wave, const float* vel, float* u2, int nx, int nz, int nb,
                      int precision) {
  return run_step(prev_wave, curr_wave, vel, u2, nx, nz, nb, true, precision);
}

int swfwi_step_undamped(float* prev_wave, const float* curr_wave, const float* vel, float* u2, int nx, int nz,
                        int precision) {
  return run_step(prev_wave, curr_wave, vel, u2, nx, nz, 0, false, precision);
}

int swfwi_expand_domain(const float* v, int nx, int nz, int nb, float dx, float dt, float* out) {
  return guarded([&] {
    NEED(v);
    NEED(out);
    if (nx <= 0 || nz <= 0 || nb < 0) fail(kErrArg, "bad model size");
    expand_domain(v, nx, nz, nb, dx, dt, out);
  });
}

int swfwi_ricker(float* wlt, int nt, float fm, float dt, float amp) {
  return guarded([&] {
    NEED(wlt);
    ricker_wavelet(wlt, nt, fm, dt, amp);
  });
}

int swfwi_random_codes(int seed, int skip, int n, int* codes) {
  return guarded([&] {
    NEED(codes);
    CodeStream cs(seed);
    if (skip > 0) cs.next(skip);
    std::vector<int> c = cs.next(n);
    for (int i = 0; i < n; i++) codes[i] = c[i];
  });
}

int swfwi_solver_create(swfwi_solver* out, int nx, int nz, int nb, float dt, float dx, float fm, int nt, int ns,
                        const int* sx, const int* sz, int ng, const int* gx, const int* gz, int precision) {
  return guarded([&] {
    NEED(out);
    NEED(sx);
    NEED(sz);
    NEED(gx);
    NEED(gz);
    if (ns <= 0 || ng <= 0) fail(kErrArg, "need sources and receivers");
    Geometry g;
    g.nx = nx;
    g.nz = nz;
    g.nb = nb;
    g.dt = dt;
    g.dx = dx;
    g.fm = fm;
    g.nt = nt;
    g.sx.assign(sx, sx + ns);
    g.sz.assign(sz, sz + ns);
    g.gx.assign(gx, gx + ng);
    g.gz.assign(gz, gz + ng);
    g.derive();  // argument errors before any device resource is created
    if (precision != SWFWI_PREC_EXACT && precision != SWFWI_PREC_FP32) fail(kErrArg, "unknown precision mode");
    std::unique_ptr<swfwi_solver_s> h(new swfwi_solver_s);
    h->s.reset(new Solver(g, precision));
    *out = h.release();
  });
}

int swfwi_solver_destroy(swfwi_solver h) {
  return guarded([&] { delete h; });
}

int swfwi_solver_dims(swfwi_solver h, int* dims) {
  return guarded([&] {
    NEED(h);
    NEED(dims);
    const Geometry& g = h->s->geo();
    dims[0] = g.nxpad;
    dims[1] = g.nzpad;
    dims[2] = g.bx0;
    dims[3] = g.bz0;
    dims[4] = g.bsize;
    dims[5] = g.ntz * g.ntx;
  });
}

int swfwi_solver_set_precision(swfwi_solver h, int precision) {
  return guarded([&] {
    NEED(h);
    if (precision != kExact && precision != kFp32) fail(kErrArg, "unknown precision mode");
    h->s->prop().set_precision(precision);
  });
}

int swfwi_solver_bind_velocity(swfwi_solver h, const float* vel_t) {
  return guarded([&] {
    NEED(h);
    NEED(vel_t);
    h->s->bind_velocity(vel_t);
  });
}

int swfwi_solver_get_velocity(swfwi_solver h, float* vel_t) {
  return guarded([&] {
    NEED(h);
    NEED(vel_t);
    h->s->download_velocity(vel_t);
  });
}

int swfwi_write_bndry(swfwi_solver h, const float* field, float* bndr) {
  return guarded([&] {
    NEED(h);
    NEED(field);
    NEED(bndr);
    Solver& s = *h->s;
    const Geometry& g = s.geo();
    s.tmp.upload(field, (size_t)s.npts(), s.stream());
    s.band.resize((size_t)g.bsize);
    op_band_pack(s.tmp.get(), s.band.get(), g.bsize, g.gridgeom(), s.stream());
    s.band.download(bndr, (size_t)g.bsize, s.stream());
    s.prop().sync();
  });
}

int swfwi_read_bndry(swfwi_solver h, const float* bndr, float* field) {
  return guarded([&] {
    NEED(h);
    NEED(field);
    NEED(bndr);
    Solver& s = *h->s;
    const Geometry& g = s.geo();
    s.tmp.upload(field, (size_t)s.npts(), s.stream());
    s.band.upload(bndr, (size_t)g.bsize, s.stream());
    op_band_unpack(s.band.get(), s.tmp.get(), g.bsize, g.gridgeom(), s.stream());
    s.tmp.download(field, (size_t)s.npts(), s.stream());
    s.prop().sync();
  });
}

int swfwi_forward_modeling(swfwi_solver h, const float* encsrc, int shot, float* dcal) {
  return guarded([&] {
    NEED(h);
    NEED(encsrc);
    NEED(dcal);
    h->s->forward_modeling(encsrc, shot, dcal);
  });
}

int swfwi_forward_modeling_noalloc(swfwi_solver h, const float* encsrc, int shot, float* dcal) {
  return guarded([&] {
    NEED(h);
    NEED(encsrc);
    NEED(dcal);
    Solver& s = *h->s;
    float* ws = noalloc_scratch(s.prop().device());
    s.forward_modeling_ws(encsrc, shot, dcal, ws, kNoAllocScratchFloats);
  });
}

int swfwi_calgradient(swfwi_solver h, const float* encsrc, int shot, const float* vsrc, float* grad) {
  return guarded([&] {
    NEED(h);
    NEED(encsrc);
    NEED(vsrc);
    NEED(grad);
    h->s->calgradient(encsrc, shot, vsrc, grad);
  });
}

int swfwi_forward_shots(swfwi_solver h, const float* wlt, float* dobs) {
  return guarded([&] {
    NEED(h);
    NEED(wlt);
    NEED(dobs);
    h->s->forward_shots(wlt, dobs);
  });
}

int swfwi_solver_counters(swfwi_solver h, double* point_updates, long long* launches) {
  return guarded([&] {
    NEED(h);
    Propagator& p = h->s->prop();
    if (point_updates) {
      point_updates[0] = p.fwd_point_updates;
      point_updates[1] = p.bwd_point_updates;
    }
    if (launches) {
      launches[0] = p.fwd_launches;
      launches[1] = p.bwd_launches;
    }
  });
}

int swfwi_solver_set_graphs(swfwi_solver h, int on) {
  return guarded([&] {
    NEED(h);
    if (on < 0 || on > 2) fail(kErrArg, "graph mode must be 0, 1 or 2");
    h->s->prop().set_graphs(on);
  });
}

int swfwi_solver_set_tuning(swfwi_solver h, int on) {
  return guarded([&] {
    NEED(h);
    h->s->prop().set_tuning(on != 0);
  });
}

int swfwi_solver_set_tile_x(swfwi_solver h, int tile_x) {
  return guarded([&] {
    NEED(h);
    h->s->prop().set_tile_x(tile_x);
  });
}

int swfwi_solver_get_tile_x(swfwi_solver h, int kind, int entries, int* tile_x) {
  return guarded([&] {
    NEED(h);
    NEED(tile_x);
    *tile_x = h->s->prop().tile_choice_cached(kind, entries);
  });
}

int swfwi_solver_time_steps(swfwi_solver h, int kind, int entries, int tile_x, int reps, float* us) {
  return guarded([&] {
    NEED(h);
    NEED(us);
    if (kind != 0 && kind != 1) fail(kErrArg, "kind must be 0 (forward) or 1 (backward)");
    if (entries < 1 || reps < 1) fail(kErrArg, "entries and reps must be >= 1");
    int width = -1;
    for (int i = 0; i < kNumWidths; i++)
      if (kTileWidths[i] == tile_x) width = i;
    if (width < 0) fail(kErrArg, "tile width must be 8, 16 or 32");
    Solver& s = *h->s;
    if (!s.rv()) fail(kErrState, "bind the velocity first");
    *us = s.prop().time_steps(kind, entries, s.rv(), 0, 1, s.w, width, reps, true);
    s.w.s_cur = s.w.s_nxt = nullptr;  // the workspace no longer holds a forward state
  });
}

int swfwi_solver_set_bwd_split(swfwi_solver h, int mode) {
  return guarded([&] {
    NEED(h);
    h->s->prop().set_bwd_split(mode);
  });
}

int swfwi_solver_set_timing(swfwi_solver h, int on) {
  return guarded([&] {
    NEED(h);
    h->s->prop().set_timing(on != 0);
  });
}

int swfwi_solver_kernel_times(swfwi_solver h, double* ms, long long* launches) {
  return guarded([&] {
    NEED(h);
    Propagator& p = h->s->prop();
    p.collect_timing();
    for (int k = 0; k < 2; k++) {
      if (ms) ms[k] = p.kernel_ms[k];
      if (launches) launches[k] = p.kernel_n[k];
    }
  });
}

int swfwi_inversion_create(swfwi_inversion* out, swfwi_solver h, const float* wlt, const float* dobs, float vmin,
                           float vmax, int nita, float maxdv) {
  return guarded([&] {
    NEED(out);
    NEED(h);
    NEED(wlt);
    NEED(dobs);
    UpdateParams p;
    p.vmin = vmin;
    p.vmax = vmax;
    p.nita = nita;
    p.maxdv = maxdv;
    std::unique_ptr<swfwi_inversion_s> inv(new swfwi_inversion_s);
    inv->inv.reset(new Inversion(*h->s, wlt, dobs, p));
    *out = inv.release();
  });
}

int swfwi_inversion_destroy(swfwi_inversion inv) {
  return guarded([&] { delete inv; });
}

int swfwi_ess_epoch(swfwi_inversion inv, int iter) {
  return guarded([&] {
    NEED(inv);
    inv->inv->ess_epoch(iter);
  });
}

int swfwi_fwi_epoch(swfwi_inversion inv, int iter) {
  return guarded([&] {
    NEED(inv);
    inv->inv->fwi_epoch(iter);
  });
}

int swfwi_inversion_stats(swfwi_inversion inv, float* out, double* ms) {
  return guarded([&] {
    NEED(inv);
    const EpochStats& s = inv->inv->stats();
    if (out) {
      out[0] = s.obj1;
      out[1] = s.steplen;
      out[2] = s.alpha2;
      out[3] = s.alpha3;
      out[4] = s.obj2;
      out[5] = s.obj3;
      out[6] = s.updateobj;
      out[7] = s.initobj;
      out[8] = (float)s.ntrial;
      out[9] = (float)s.parabolic;
    }
    if (ms) {
      ms[0] = s.grad_ms;
      ms[1] = s.linesearch_ms;
    }
  });
}

int swfwi_inversion_gradient(swfwi_inversion inv, float* grad) {
  return guarded([&] {
    NEED(inv);
    NEED(grad);
    inv->inv->download_grad(grad);
  });
}

int swfwi_inversion_direction(swfwi_inversion inv, float* dir) {
  return guarded([&] {
    NEED(inv);
    NEED(dir);
    inv->inv->download_direction(dir);
  });
}

}  // extern "C"

// ---- ensemble Kalman analysis ------------------------------------------------
#include "enkf.h"

struct swfwi_enkf_s {
  std::unique_ptr<swfwi::Enkf> e;
};

extern "C" {

int swfwi_enkf_create(swfwi_enkf* out, swfwi_solver h, const float* wlt, const float* dobs, float sigfac,
                      int nmembers) {
  return guarded([&] {
    NEED(out);
    NEED(h);
    NEED(wlt);
    NEED(dobs);
    std::unique_ptr<swfwi_enkf_s> e(new swfwi_enkf_s);
    e->e.reset(new Enkf(*h->s, wlt, dobs, sigfac, nmembers));
    *out = e.release();
  });
}

int swfwi_enkf_destroy(swfwi_enkf e) {
  return guarded([&] { delete e; });
}

int swfwi_enkf_set_members(swfwi_enkf e, const float* vels) {
  return guarded([&] {
    NEED(e);
    NEED(vels);
    e->e->set_members(vels);
  });
}

int swfwi_enkf_get_members(swfwi_enkf e, float* vels) {
  return guarded([&] {
    NEED(e);
    NEED(vels);
    e->e->get_members(vels);
  });
}

int swfwi_enkf_analyze(swfwi_enkf e, const double* perturbation, unsigned long long seed) {
  return guarded([&] {
    NEED(e);
    e->e->analyze(perturbation, seed);
  });
}

int swfwi_enkf_mean(swfwi_enkf e, float* mean) {
  return guarded([&] {
    NEED(e);
    NEED(mean);
    e->e->mean(mean);
  });
}

int swfwi_enkf_info(swfwi_enkf e, double* sigma, int* clip, double* svals, double* gain) {
  return guarded([&] {
    NEED(e);
    if (sigma) *sigma = e->e->sigma();
    if (clip) *clip = e->e->clip();
    const int N = e->e->members();
    if (svals)
      for (int k = 0; k < N && k < (int)e->e->svals().size(); k++) svals[k] = e->e->svals()[k];
    if (gain)
      for (size_t k = 0; k < e->e->gain().size(); k++) gain[k] = e->e->gain()[k];
  });
}

}  // extern "C"

// ---- ensemble-Kalman FWI -------------------------------------------------------
#include "ensemble.h"

struct swfwi_enfwi_s {
  std::unique_ptr<swfwi::EnsembleFwi> f;
};

extern "C" {

int swfwi_enfwi_create(swfwi_enfwi* out, swfwi_solver h, swfwi_enkf e, float vmin, float vmax, int nita,
                       float maxdv) {
  return guarded([&] {
    NEED(out);
    NEED(h);
    NEED(e);
    UpdateParams p;
    p.vmin = vmin;
    p.vmax = vmax;
    p.nita = nita;
    p.maxdv = maxdv;
    std::unique_ptr<swfwi_enfwi_s> f(new swfwi_enfwi_s);
    f->f.reset(new EnsembleFwi(*h->s, *e->e, p));
    *out = f.release();
  });
}

int swfwi_enfwi_destroy(swfwi_enfwi f) {
  return guarded([&] { delete f; });
}

int swfwi_enfwi_epoch(swfwi_enfwi f, int iter) {
  return guarded([&] {
    NEED(f);
    if (iter < 0) fail(kErrArg, "iteration must be >= 0");
    f->f->epoch(iter);
  });
}

int swfwi_enfwi_mean_objective(swfwi_enfwi f, float* mean, float* obj) {
  return guarded([&] {
    NEED(f);
    NEED(obj);
    *obj = f->f->mean_objective(mean);
  });
}

int swfwi_enfwi_stats(swfwi_enfwi f, float* out, double* ms, int* counts) {
  return guarded([&] {
    NEED(f);
    const std::vector<EpochStats>& st = f->f->stats();
    if (out)
      for (size_t i = 0; i < st.size(); i++) {
        const EpochStats& s = st[i];
        const float v[10] = {s.obj1, s.steplen, s.alpha2, s.alpha3, s.obj2, s.obj3, s.updateobj, s.initobj,
                             (float)s.ntrial, (float)s.parabolic};
        for (int k = 0; k < 10; k++) out[i * 10 + k] = v[k];
      }
    if (ms) {
      ms[0] = f->f->grad_ms();
      ms[1] = f->f->linesearch_ms();
    }
    if (counts) {
      counts[0] = f->f->trial_rounds();
      counts[1] = f->f->trials();
    }
  });
}

int swfwi_enfwi_gradients(swfwi_enfwi f, float* grads) {
  return guarded([&] {
    NEED(f);
    NEED(grads);
    f->f->download_grads(grads);
  });
}

// ---- mixed-size batch with reduced-precision checkpoints ------------------

int swfwi_batch_create(swfwi_batch* out, int ngroups, const int* dims, const float* params, const int* sx,
                       const int* sz, int snap_prec, int order) {
  return guarded([&] {
    NEED(out);
    NEED(dims);
    NEED(params);
    NEED(sx);
    NEED(sz);
    if (ngroups < 1) fail(kErrArg, "a batch needs at least one group");
    std::vector<BatchGroup> gs((size_t)ngroups);
    int o = 0;
    for (int g = 0; g < ngroups; g++) {
      const int* d = dims + 9 * g;
      BatchGroup& G = gs[g];
      G.nx = d[0];
      G.nz = d[1];
      G.nb = d[2];
      G.nt = d[3];
      const int ns = d[4];
      G.gx0 = d[5];
      G.ng = d[6];
      G.gz = d[7];
      G.nmodels = d[8];
      G.dx = params[4 * g];
      G.dt = params[4 * g + 1];
      G.fm = params[4 * g + 2];
      G.dscale = params[4 * g + 3];
      if (ns < 1) fail(kErrArg, "a group needs at least one shot");
      G.sx.assign(sx + o, sx + o + ns);
      G.sz.assign(sz + o, sz + o + ns);
      o += ns;
    }
    *out = nullptr;
    std::unique_ptr<swfwi_batch_s> h(new swfwi_batch_s);
    h->b.reset(new Batch(gs, snap_prec, order));
    *out = h.release();
  });
}

int swfwi_batch_destroy(swfwi_batch b) {
  return guarded([&] { delete b; });
}

int swfwi_batch_set_snapshot_precision(swfwi_batch b, int snap_prec) {
  return guarded([&] {
    NEED(b);
    b->b->set_snap_prec(snap_prec);
  });
}

int swfwi_batch_set_graphs(swfwi_batch b, int on) {
  return guarded([&] {
    NEED(b);
    b->b->set_graphs(on != 0);
  });
}

int swfwi_batch_set_tuning(swfwi_batch b, int on) {
  return guarded([&] {
    NEED(b);
    b->b->set_tuning(on != 0);
  });
}

int swfwi_batch_set_tile_x(swfwi_batch b, int tile_x) {
  return guarded([&] {
    NEED(b);
    b->b->set_tile_x(tile_x);
  });
}

int swfwi_batch_set_block_order(swfwi_batch b, int order) {
  return guarded([&] {
    NEED(b);
    b->b->set_block_order(order);
  });
}

int swfwi_batch_block_order(swfwi_batch b, int* order) {
  return guarded([&] {
    NEED(b);
    NEED(order);
    *order = b->b->block_order();
  });
}

int swfwi_batch_tile_x(swfwi_batch b, int* tile_x) {
  return guarded([&] {
    NEED(b);
    NEED(tile_x);
    *tile_x = b->b->tile_x();
  });
}

int swfwi_batch_bind_model(swfwi_batch b, int group, int model, const float* vel_t) {
  return guarded([&] {
    NEED(b);
    NEED(vel_t);
    b->b->bind_model(group, model, vel_t);
  });
}

int swfwi_batch_forward(swfwi_batch b, float* dcal) {
  return guarded([&] {
    NEED(b);
    b->b->forward(dcal);
  });
}

int swfwi_batch_adjoint_source(swfwi_batch b, const float* dobs, float* vsrc) {
  return guarded([&] {
    NEED(b);
    NEED(dobs);
    b->b->adjoint_from_observed(dobs, vsrc);
  });
}

int swfwi_batch_backward(swfwi_batch b, const float* vsrc, float* grads, float* model_grads) {
  return guarded([&] {
    NEED(b);
    b->b->backward(vsrc, grads, model_grads);
  });
}

int swfwi_batch_info(swfwi_batch b, double* out) {
  return guarded([&] {
    NEED(b);
    NEED(out);
    out[0] = b->b->entries();
    out[1] = b->b->forward_points();
    out[2] = b->b->backward_points();
    out[3] = (double)b->b->snapshot_bytes();
    out[4] = b->b->last_ms[0];
    out[5] = b->b->last_ms[1];
  });
}

int swfwi_batch_set_taps(swfwi_batch b, int n, const int* steps, int oz) {
  return guarded([&] {
    NEED(b);
    if (n < 0 || (n > 0 && !steps)) fail(kErrArg, "bad tap list");
    b->b->set_taps(n > 0 ? std::vector<int>(steps, steps + n) : std::vector<int>(), oz);
  });
}

int swfwi_batch_taps(swfwi_batch b, float* out) {
  return guarded([&] {
    NEED(b);
    NEED(out);
    b->b->download_taps(out);
  });
}

int swfwi_batch_step_from(swfwi_batch b, const float* prev, const float* cur, int it0, int nsteps, float* prev_out,
                          float* cur_out) {
  return guarded([&] {
    NEED(b);
    NEED(prev);
    NEED(cur);
    b->b->step_from(prev, cur, it0, nsteps, prev_out, cur_out);
  });
}

int swfwi_batch_set_correction(swfwi_batch b, swfwi_net net) {
  return guarded([&] {
    NEED(b);
    b->b->set_correction(net ? net->n.get() : nullptr);
  });
}

// ---- learned coarse-grid correction network --------------------------------

namespace {
NetGeom net_geom(int nx, int nz, int entries, const int* region) {
  NEED(region);
  NetGeom g{nx, nz, (long long)nx * nz, entries, region[0], region[1], region[2], region[3], nullptr, 0, 1};
  g.rstride = g.stride;
  if (nx < 1 || nz < 1 || entries < 1) fail(kErrArg, "bad field stack");
  return g;
}
}  // namespace

int swfwi_net_create(swfwi_net* out) {
  return guarded([&] {
    NEED(out);
    *out = nullptr;
    std::unique_ptr<swfwi_net_s> h(new swfwi_net_s);
    h->n.reset(new ConvNet());
    h->s.reset(new Stream());
    *out = h.release();
  });
}

int swfwi_net_destroy(swfwi_net n) {
  return guarded([&] { delete n; });
}

int swfwi_net_set_weights(swfwi_net n, const float* w) {
  return guarded([&] {
    NEED(n);
    NEED(w);
    n->n->set_weights(w);
  });
}

int swfwi_net_get_weights(swfwi_net n, float* w) {
  return guarded([&] {
    NEED(n);
    NEED(w);
    n->n->get_weights(w);
  });
}

int swfwi_net_apply(swfwi_net n, const float* fields, const float* rv, int nx, int nz, int entries,
                    const int* region, float* y) {
  return guarded([&] {
    NEED(n);
    NEED(fields);
    NEED(rv);
    NEED(y);
    NetGeom g = net_geom(nx, nz, entries, region);
    hipStream_t s = n->s->get();
    DBuf<float> p, r, out;
    p.upload(fields, (size_t)g.stride * entries, s);
    r.upload(rv, (size_t)g.stride * entries, s);
    g.r = r.get();
    out.resize((size_t)g.stride * entries);
    launch_fill(out.get(), g.stride * entries, 0.f, s);
    net_apply(p.get(), out.get(), g, n->n->dev(), s);
    out.download(y, (size_t)g.stride * entries, s);
    SWFWI_HIP(hipStreamSynchronize(s));
  });
}

int swfwi_net_time_apply(swfwi_net n, const float* fields, const float* rv, int nx, int nz, int entries,
                         const int* region, int reps, double* ms) {
  return guarded([&] {
    NEED(n);
    NEED(fields);
    NEED(rv);
    NEED(ms);
    if (reps < 1) fail(kErrArg, "reps must be >= 1");
    NetGeom g = net_geom(nx, nz, entries, region);
    hipStream_t s = n->s->get();
    DBuf<float> p, r, out;
    p.upload(fields, (size_t)g.stride * entries, s);
    r.upload(rv, (size_t)g.stride * entries, s);
    g.r = r.get();
    out.resize((size_t)g.stride * entries);
    net_apply(p.get(), out.get(), g, n->n->dev(), s);  // warm
    hipEvent_t e0, e1;
    SWFWI_HIP(hipEventCreate(&e0));
    SWFWI_HIP(hipEventCreate(&e1));
    SWFWI_HIP(hipEventRecord(e0, s));
    for (int k = 0; k < reps; k++) net_apply(p.get(), out.get(), g, n->n->dev(), s);
    SWFWI_HIP(hipEventRecord(e1, s));
    SWFWI_HIP(hipEventSynchronize(e1));
    float t = 0;
    SWFWI_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t / reps;
  });
}

int swfwi_net_loss_grad(swfwi_net n, const float* in, const float* rv, const float* target, int nx, int nz,
                        int entries, const int* region, float* grad, double* loss) {
  return guarded([&] {
    NEED(n);
    NEED(in);
    NEED(rv);
    NEED(target);
    NetGeom g = net_geom(nx, nz, entries, region);
    hipStream_t s = n->s->get();
    DBuf<float> p, r, t;
    p.upload(in, (size_t)g.stride * entries, s);
    r.upload(rv, (size_t)g.stride * entries, s);
    t.upload(target, (size_t)g.stride * entries, s);
    g.r = r.get();
    const double l = n->n->loss_grad(p.get(), t.get(), g, grad, s);
    if (loss) *loss = l;
  });
}

int swfwi_net_set_data(swfwi_net n, const float* in, const float* rv, const float* target, int nx, int nz,
                       int nsamples, const int* region) {
  return guarded([&] {
    NEED(n);
    NEED(in);
    NEED(rv);
    NEED(target);
    n->n->set_data(in, rv, target, nsamples, net_geom(nx, nz, nsamples, region));
  });
}

int swfwi_net_train(swfwi_net n, int batch, int steps, float lr, unsigned long long seed, double* losses) {
  return guarded([&] {
    NEED(n);
    n->n->train(batch, steps, lr, seed, losses, n->s->get());
  });
}

}  // extern "C"
