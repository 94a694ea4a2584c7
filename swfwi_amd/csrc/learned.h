// Learned coarse-grid correction: a 3-layer 3x3 convolution network
// (2 -> 16 -> 16 -> 1 channels, ReLU between, no biases) that maps a
// coarse-grid pressure field p to a correction y, applied as p <- p + y.
// The two input channels are p and r p, r = 1 / vel_t the local Courant
// number squared: both scale with p, so the bias-free ReLU network is
// positively homogeneous in the wavefield as the wave equation is linear,
// and r tells it the local velocity the stencil's dispersion depends on.
// Every layer runs on BF16 MFMA with FP32 accumulation
// (v_mfma_f32_16x16x16_bf16): the inputs p, r p, the hidden activations and
// the weights enter the products rounded to BF16, sums and the output y are
// FP32.  Training kernels form the loss gradient with respect to every
// weight (the 16 -> 16 products again on BF16 MFMA) and update the FP32
// master weights with Adam.
//
// Weights, flattened (kNetParams floats): W1[t][k][c] (9 x 2 x 16, k = 0
// for p, 1 for r p), W2[t][ci][co] (9 x 16 x 16), W3[t][ci] (9 x 16), tap
// t = (dx + 1) * 3 + (dz + 1) for the neighbour at (x + dx, z + dz); the
// fields are zero outside their array.
#pragma once
#include <cstdint>
#include <vector>

#include "device.h"

namespace swfwi {

constexpr int kNetC = 16;
constexpr int kNetIn = 2;
constexpr int kNetW1 = 9 * kNetIn * kNetC;
constexpr int kNetW2 = 9 * kNetC * kNetC;
constexpr int kNetW3 = 9 * kNetC;
constexpr int kNetParams = kNetW1 + kNetW2 + kNetW3;

// A stack of `entries` fields of one padded grid (nx x nz, z contiguous,
// entry e at base + e * stride); the network's outputs cover the region
// [x0, x1) x [z0, z1).
struct NetGeom {
  int nx, nz;
  long long stride;
  int entries;
  int x0, x1, z0, z1;
  const float* r;     // 1 / vel_t of entry e at r + (e / rdiv) * rstride, same layout
  long long rstride;
  int rdiv;
};

// Device weights in the layouts the kernels read: every layer's BF16 MFMA
// operands packed per lane (W2 also transposed for the backward pass) and
// W3 rounded to BF16, as FP32.
struct NetDev {
  const uint16_t* w1a;  // layer-1 A operands, [2][64 lanes][4]
  const uint16_t* w2b;  // backward B operands, [9][64][4] (co = 4 (l / 16) + v, ci = l % 16)
  const uint16_t* w3a;  // layer-3 A operands, [64][4]
  const uint16_t* w2p;  // layer-2 A operands of the K = 32 products (taps 2q, 2q+1; tap 8, zeros), [5][64][8]
  const float* w3;      // W3 rounded to BF16
};

void net_pack(const float* theta, const NetDev& w, hipStream_t s);
// y = net(p) over the region (y elsewhere untouched)
void net_apply(const float* p, float* y, const NetGeom& g, const NetDev& w, hipStream_t s);
// p += y over the region
void net_add(float* p, const float* y, const NetGeom& g, hipStream_t s);
// Loss of a batch: r = p + y - t over the region, dy = 2 r / count, the sum
// of r^2 (FP64) into *loss (device), dy zero outside the region.
void net_loss(const float* p, const float* y, const float* t, float* dy, const NetGeom& g, double* loss_parts,
              int nparts, hipStream_t s);
// d loss / d weights: per-block partials reduced in a fixed order into grad.
// Returns nothing; `parts` must hold net_backward_blocks(g) * kNetParams floats.
long long net_backward_blocks(const NetGeom& g);
void net_backward(const float* p, const float* dy, const NetGeom& g, const NetDev& w, float* parts, float* grad,
                  hipStream_t s);
// Adam (step >= 1): theta -= lr * mhat / (sqrt(vhat) + eps)
void net_adam(float* theta, const float* grad, float* m, float* v, int step, float lr, float b1, float b2, float eps,
              hipStream_t s);
void net_reduce_parts(const double* parts, int nparts, double* out, hipStream_t s);

// Host handle: weights (FP32 master copy + packed device forms), Adam state
// and scratch for training on stacks of coarse fields.
class ConvNet {
 public:
  ConvNet();
  void set_weights(const float* host);
  void get_weights(float* host) const;
  NetDev dev() const { return NetDev{w1a_.get(), w2b_.get(), w3a_.get(), w2p_.get(), w3_.get()}; }
  // p <- p + net(p) over the region, for a stack of fields (device)
  void correct(float* p, const NetGeom& g, float* yscratch, hipStream_t s) const;
  // One Adam step on the batch (inputs, targets on the device); the loss
  // sum_r r^2 before the step into *loss_dev (device, FP64).
  void train_step(const float* in, const float* target, const NetGeom& g, float lr, double* loss_dev,
                  hipStream_t s);
  // loss sum into *loss_dev and the weight gradient, without a wait
  void loss_grad_async(const float* in, const float* target, const NetGeom& g, double* loss_dev, hipStream_t s);
  // loss and gradient without an update (for checks)
  double loss_grad(const float* in, const float* target, const NetGeom& g, float* grad_host, hipStream_t s);
  int steps() const { return step_; }
  // Training set on the device: inputs / targets [n][nx][nz] (one padded
  // coarse grid), the loss region in every sample.
  void set_data(const float* in_host, const float* r_host, const float* target_host, int n, const NetGeom& g);
  // `steps` Adam steps over minibatches of `batch` samples drawn in a
  // seeded shuffled order; losses[k] = the loss before step k.
  void train(int batch, int steps, float lr, unsigned long long seed, double* losses, hipStream_t s);

 private:
  void repack(hipStream_t s);
  DBuf<float> theta_, m_, v_, grad_, w3_;
  DBuf<uint16_t> w1a_, w2b_, w3a_, w2p_;
  DBuf<float> y_, dy_, parts_, din_, dr_, dtgt_, bin_, br_, btgt_;
  NetGeom dg_{};
  int dn_ = 0;
  DBuf<double> lparts_, lsum_;
  int step_ = 0;
};

}  // namespace swfwi
