// net.h -- the learned preconditioner network on the GPU (host interface).
//
// A 4-level U-Net on the cell grid (no reference counterpart), channels
// (16, 32, 64, 64): encoder E0 1->16, then per level a 2x2 average pool and
// E1 16->32, E2 32->64, E3 64->64 (3x3 conv + ReLU); decoder D3 64->64 on
// the top encoder output, then per level a nearest 2x upsample of the
// coarser output plus the encoder output, D2 64->32, D1 32->16 (+ ReLU),
// D0 16->1.  Convolutions are 'same' 3x3 with zero padding and no bias, so
// the network is positively homogeneous.  The 16..64-channel convolutions
// (forward, data gradient and weight gradient) run as BF16 implicit GEMMs on
// gfx950's MFMA v_mfma_f32_16x16x32_bf16 with FP32 accumulation;
// activations are BF16 NHWC, master weights FP32 (Adam).
// Weight packing: per layer W[tap][ci][co], tap = 3*ky+kx, offsets
// (kx-1, ky-1), layers in the order E0..E3, D3..D0.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace hg {

class UNet {
 public:
  static constexpr int kLevels = 4;
  UNet(int nx, int ny);
  ~UNet();
  UNet(const UNet&) = delete;
  UNet& operator=(const UNet&) = delete;
  static bool Supported(int nx, int ny) { return nx % 8 == 0 && ny % 8 == 0 && nx >= 16 && ny >= 16; }
  int num_weights() const { return nw_; }

  void SetWeights(const float* host, hipStream_t st);   // FP32 master weights
  void GetWeights(float* host, hipStream_t st);
  void InitWeights(uint64_t seed, hipStream_t st);
  // y = net(x) on level 0 (FP32 in/out, device); with `base`, y = base + net(x).
  void Forward(const float* x, float* y, const float* base, hipStream_t st, const int* done);
  // Gradients of sum(gy * net(x)) into the FP32 weight-gradient buffer
  // (accumulated; Forward(x) must have run last).  gx may be null.
  void Backward(const float* x, const float* gy, hipStream_t st);
  void ZeroGrad(hipStream_t st);
  void GetGrad(float* host, hipStream_t st);
  // One Adam step on the accumulated gradient.
  void AdamStep(float lr, hipStream_t st);
  // Measures the training loss in the energy norm of a face-form operator
  // (FP32 device coefficients kx[(nx+1)*ny], ky[nx*(ny+1)], kept by
  // pointer): loss = d^T A d / n with d = net(x) - target.  Null: mean
  // squares.
  // Without activations the network is a linear operator of its input
  // (a learned linear multigrid-like map); default: ReLU after every layer
  // but the last.
  void SetLinear(bool linear) { linear_ = linear || tied_; }
  bool linear() const { return linear_; }
  // Tied: decoder layer D_l is the transpose of encoder layer E_l (taps
  // flipped, channels swapped) and there are no activations.  Nearest
  // upsampling is 4x the adjoint of 2x2 average pooling, so the network is
  // then y = sum_l 4^l F_l^T F_l x with F_l = E_l P ... P E_0: a symmetric
  // positive semi-definite operator, an additive multilevel (BPX-like)
  // preconditioner with learned stencils.  The decoder weight slots are
  // unused (their gradient stays zero).
  void SetTied(bool tied) {
    tied_ = tied;
    if (tied) linear_ = true;
  }
  bool tied() const { return tied_; }
  // Scale-free energy loss (with a metric): loss = 1 - cos^2 of o*net(x) and
  // the target in the energy inner product, i.e. the relative energy error
  // of the best multiple alpha*o*net(x); last_alpha() is that multiple for
  // the last training step.
  void SetScaleFree(bool on) { scale_free_ = on; }
  double last_alpha() const { return last_alpha_; }
  // Multiplies the network's output by a > 0 (folded into the weights).
  void ScaleOutput(float a, hipStream_t st);
  // Per-cell factor o applied to the output inside the training loss
  // (loss of o*net(x) - target); null: none.  Kept by pointer.
  void SetOutputScale(const float* o) { oscale_ = o; }
  void SetMetric(const float* kx, const float* ky) {
    mkx_ = kx;
    mky_ = ky;
  }
  // loss = mean((net(x) - target)^2) (or the metric above): forward, loss gradient, backward and
  // one Adam step.  With `loss` set the step waits for the loss value.
  void TrainStep(const float* x, const float* target, float lr, hipStream_t st, double* loss);
  // Arithmetic of one forward pass and of one training step (multiply-adds
  // counted as two), for throughput figures.
  double ForwardFlops() const;
  double TrainFlops() const;
  // FLOPs of the layers that run on MFMA (the 16..64-channel convolutions).
  double MfmaForwardFlops() const;

 private:
  int nx_, ny_, nw_;
  int lw_[kLevels], lh_[kLevels];
  int ch_[kLevels];
  // layer table: offset into the packed weights, cin, cout, level
  struct Layer {
    int off, cin, cout, level;
  };
  Layer layer_[2 * kLevels];
  float *w_ = nullptr, *g_ = nullptr, *m_ = nullptr, *v_ = nullptr;  // FP32 master / grad / Adam
  uint16_t *wf_ = nullptr, *wb_ = nullptr;  // BF16 weights: forward [tap][co][ci], dgrad flipped
  int adam_t_ = 0;
  // activations (BF16 NHWC) per level
  uint16_t *enc_[kLevels] = {}, *pin_[kLevels] = {}, *din_[kLevels] = {}, *dout_[kLevels] = {};
  // gradients (BF16 NHWC)
  uint16_t *genc_[kLevels] = {}, *gtmp_[kLevels] = {}, *gpre_[kLevels] = {}, *gin_[kLevels] = {};
  float *yb_ = nullptr, *gy_ = nullptr, *qu_ = nullptr, *qt_ = nullptr;
  bool scale_free_ = false;
  double last_alpha_ = 0.;
  double* part_ = nullptr;
  double* loss_h_ = nullptr;
  const float *mkx_ = nullptr, *mky_ = nullptr;
  bool linear_ = false, tied_ = false;
  const float* oscale_ = nullptr;
  float* wd0_ = nullptr;   // tied first-layer transpose (FP32 [tap][ci])
  float* gtie_ = nullptr;  // weight-gradient scratch of a tied decoder layer
  const float* D0Weights() const;
  void ConvertWeights(hipStream_t st);
};

}  // namespace hg
