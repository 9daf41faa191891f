// learned_net.h -- the learned preconditioner of the flexible CG solver.
//
// M(r) = V(r) + N(D^-1 (r - A V(r))): one multigrid V-cycle V (mg.h), then
// the U-Net N (net.h) maps the Jacobi-scaled residual left by the V-cycle
// to a correction of the V-cycle's output.  With N's last layer at zero the
// preconditioner is the V-cycle itself, so an untrained or badly trained net
// never leaves the multigrid baseline by much, and flexible CG (with its
// Jacobi-PCG fallback) keeps the solve correct whatever N returns.
//
// Training pairs come from the solver on the operator in use: for a set of
// right-hand sides b, x* = A^-1 b from a tight FCG + V-cycle solve, then the
// stationary iteration x_{k+1} = x_k + V(b - A x_k) yields residuals r_k and
// errors e_k = x* - x_k; the pair is (input D^-1 (r_k - A V(r_k)), target
// e_k - V(r_k)), both scaled by 1/rms(target) (the net is positively
// homogeneous, so the scaling changes nothing at inference) and the target
// projected to zero mean when the operator is singular.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>

#include "hydro_gpu.h"
#include "learned.h"
#include "mg.h"
#include "net.h"

namespace hg {

class CgSolver;

struct TrainOpts {
  int rhs_count = 4;     // random multi-scale right-hand sides
  int mg_iters = 12;     // stationary V-cycle iterations per right-hand side (pairs)
  int steps = 300;       // Adam steps (one pair each)
  float lr = 1e-3f;      // peak learning rate (cosine decay to lr/10)
  uint64_t seed = 1;
  bool project = true;   // operator singular (constants in the null space)
  bool linear = false;      // network without activations (a linear operator)
  bool tied = true;         // symmetric network, M = [V +] S N S with S = D^-1/2
  bool standalone = false;  // no V-cycle: M = N(D^-1 r) (or S N S when tied)
};

// Options of the C-ABI (hg_train_opts) checked and converted; throws
// std::invalid_argument on values out of range.
TrainOpts TrainOptsFromAbi(const hg_train_opts& o, bool project);

struct TrainReport {
  int pairs = 0;
  int steps = 0;
  double loss_first = 0.;  // mean relative loss of the first 10% of steps
  double loss_last = 0.;   // ... and of the last 10%
  double alpha = 0.;       // output scale folded into the net after training
  // held-out acceptance solve (relative residual 1e-8): iterations and device
  // ms with the learned preconditioner and with the method without it
  int val_iters = 0, val_iters_baseline = 0;
  double val_ms = 0., val_ms_baseline = 0.;
  bool accepted = false;   // the learned preconditioner is used from now on
  double ms = 0.;          // wall time of pair generation + training
};

class LearnedNet : public Precond {
 public:
  // mg must stay alive and be built on the operator before Apply/Train.
  LearnedNet(int nx, int ny, MgPrecond* mg);
  ~LearnedNet() override;
  LearnedNet(const LearnedNet&) = delete;
  LearnedNet& operator=(const LearnedNet&) = delete;
  static bool Supported(int nx, int ny) { return UNet::Supported(nx, ny); }
  // True once training produced a net that beat the method without it on
  // the held-out acceptance solve (fewer iterations and less device time);
  // the solvers use the net only then.
  bool Trained() const { return trained_; }
  void Apply(const double* r, float* z, hipStream_t st, const int* d_done) override;
  // Generates pairs on cg's operator (cg.SetOperator and mg->Build done) and
  // trains the net on them, on cg's stream.
  TrainReport Train(CgSolver& cg, const TrainOpts& o);
  UNet& net() { return *net_; }

 private:
  int nx_, ny_;
  MgPrecond* mg_;
  std::unique_ptr<UNet> net_;
  bool trained_ = false;
  bool standalone_ = false;
  float* s_ = nullptr;    // net input at Apply
  float* y_ = nullptr;    // net output at Apply (tied)
  float* dsq_ = nullptr;  // D^-1/2 of the finest level, for the training loss
};

}  // namespace hg
