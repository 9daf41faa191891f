// learned.h -- interface of the learned preconditioner used by flexible CG.
#pragma once

#include <hip/hip_runtime.h>

namespace hg {

// z = M(r) for FP32 vectors on the cell grid.  Kernels must return at once
// when *d_done is non-zero (the solve already converged).
class LearnedPrecond {
 public:
  virtual ~LearnedPrecond() {}
  virtual void Apply(const float* r, float* z, hipStream_t st, const int* d_done) = 0;
};

}  // namespace hg
