// learned_net.h -- the learned multigrid-like preconditioner (host interface).
#pragma once

#include <hip/hip_runtime.h>

#include "learned.h"

namespace hg {

class LearnedNet : public Precond {
 public:
  LearnedNet(int nx, int ny, int levels, int channels, hipStream_t st);
  ~LearnedNet() override;
  bool Trained() const { return trained_; }
  void Apply(const double* r, float* z, hipStream_t st, const int* d_done) override;

 private:
  int nx_, ny_, levels_, channels_;
  hipStream_t st_;
  bool trained_ = false;
};

}  // namespace hg
