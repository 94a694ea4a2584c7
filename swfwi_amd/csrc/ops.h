// Launchers of the data-side kernels around the propagator: encoding, mute,
// residual/objective, time derivative, gradient post-processing, conjugate
// gradient, step-length bound, model update and the FP64 EnKF products.
// Every reduction accumulates in FP64 in a fixed order (deterministic), with
// per-block partials reduced by wavefront shuffles.
#pragma once
#include <hip/hip_runtime.h>

namespace swfwi {

struct GridGeom {  // padded geometry, as in kernels.h Grid but host-side only
  int nx, nz, ld, bx0, bxn, bz0, bzn;
};

// encObs[j] = sum_s dobs[s][j] * code[s], sequential over s in FP32
// (Encoder::encodeObsData, encoder.cpp:29-42).
void op_encode_obs(const float* dobs, const int* code, int ns, long long m, float* out, hipStream_t s);
// encSrc[it][s] = wlt[it] * code[s] (Encoder::encodeSource, encoder.cpp:14-27).
void op_encode_source(const float* wlt, const int* code, int ns, int nt, float* out, hipStream_t s);

// Mute windows of removeDirectArrival (damp4t10d.cpp:321-365) for E models:
// vel + e*vstride; source position (sx, sz) padded coords per entry; receivers
// gx/gz padded coords.  start/end are [E][ng].
void op_mute_windows(const float* vel, long long vstride, int E, const GridGeom& g, const int* sxp, const int* szp,
                     const int* gxp, const int* gzp, int ng, int nt, int half_len, int* start, int* end,
                     hipStream_t s);
// Zero data[e][t][r] for t in [start[e*wstride + r], end[...]).
void op_mute_apply(float* data, long long dstride, int E, const int* start, const int* end, long long wstride,
                   int nt, int ng, hipStream_t s);

// res[e] = obs(+e*ostride) - cal[e]; partial FP64 sums of res^2 into
// partials[E][nblk]; returns nblk (per entry).
int op_residual(const float* obs, long long ostride, const float* cal, float* res, long long m, int E,
                double* partials, hipStream_t s);
// Sum of squares (FP64) of x[e] over m, partials as above.
int op_sumsq(const float* x, long long m, int E, double* partials, hipStream_t s);
// out[e] = sum of partials[e][0..nblk) in a fixed order.
void op_reduce_partials(const double* partials, int nblk, int E, double* out, hipStream_t s);

// second_order_virtual_source_forth_accuracy over every trace
// (fwiframework.cpp:75-105): out[e][t][r] from in[e][t-2..t+2][r] in FP64.
void op_trans_vsrc(const float* in, float* out, int nt, int ng, int E, hipStream_t s);

// scaleGradient then maskGradient (damp4t10d.cpp:309-319,216-226) on E grads.
void op_scale_mask(float* grad, long long gstride, int E, const GridGeom& g, hipStream_t s);
// out = sum over e of in[e] (FP32, ascending e, starting from 0).
void op_sum_entries(const float* in, long long stride, int E, long long n, float* out, hipStream_t s);

// FP64 partial dots for updateGrad (fwiframework.cpp:42-72): a=sum c*c,
// b=sum c*p, c=sum p*p; partials [3][nblk].
int op_cg_dots(const float* cur, const float* pre, long long n, double* partials, hipStream_t s);
// dir = cur + beta*dir; pre = cur.
void op_cg_update(const float* cur, float* pre, float* dir, float beta, long long n, hipStream_t s);

// calMaxAlpha2_3 (updatesteplenop.cpp:25-47): min over points of
// (tmpv - vel)/|grad|; writes the float result to out[0] (FLT_MAX if none).
void op_max_alpha(const float* vel, const float* grad, long long n, float dt, float dx, float maxdv, float* scratch,
                  float* out, hipStream_t s);

// update_vel (updatevelop.cpp:11-21): nv[e] = clamp(v + steps[e]*dir) for E
// step lengths (host array copied by value through a small device table).
void op_update_vel(const float* v, const float* dir, const float* steps, int E, long long n, float tmin, float tmax,
                   float* nv, long long nvstride, hipStream_t s);
// refillBoundary (damp4t10d.cpp:228-252), in place, E models.
void op_refill(float* a, long long stride, int E, const GridGeom& g, hipStream_t s);

// writeBndry / readBndry of one step (damp4t10d.cpp:478-536): the 6-wide
// band of a padded field to band[bsize] and back.
void op_band_pack(const float* field, float* band, int bsize, const GridGeom& g, hipStream_t s);
void op_band_unpack(const float* band, float* field, int bsize, const GridGeom& g, hipStream_t s);

// rv = 1/vel for E models.
void op_reciprocal(const float* vel, float* rv, long long n, hipStream_t s);

// Device-to-device copy bandwidth (GB/s, read + write bytes) of a plain
// float4 grid-stride copy of n floats, `iters` back-to-back launches.
double op_copy_bandwidth(long long n, int iters, hipStream_t s);

// Microseconds per launch of `reps` back-to-back kernels of `blocks` blocks
// that each write one float (from a captured graph when `graph`).
double op_launch_latency(int blocks, int reps, bool graph, hipStream_t s);

// Add `src` (+e*stride) into dst (+e*stride) elementwise (FP32).
void op_axpy1(const float* src, float* dst, long long n, hipStream_t s);

// ---- EnKF (FP64) -------------------------------------------------------------
// C[N][N] (column-major, C[i + N*j]) += sum_r A[i][r] * B[j][r] over rows r,
// where A and B are E x m matrices stored entry-major (A[i][r] = A + i*m + r,
// FP32 or FP64 inputs promoted to FP64).  Used for the data-space Gram
// products of the analysis (enkfanalyze.cpp:420-476).
void op_gram_f64(const double* A, const double* B, int N, long long m, double* C, hipStream_t s);
// model update t5 = A' * T (A' = members - mean, modelSize x N, T N x N).
void op_apply_nxn(const double* Ap, const double* T, int N, long long n, double* out, hipStream_t s);

}  // namespace swfwi
