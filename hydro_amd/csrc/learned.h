// learned.h -- preconditioner interface of the flexible CG solver.
#pragma once

#include <hip/hip_runtime.h>

namespace hg {

// z = M(r): r is the FP64 residual on the cell grid, z the FP32 correction.
// Kernels must return at once when *d_done is non-zero (solve converged).
class Precond {
 public:
  virtual ~Precond() {}
  virtual void Apply(const double* r, float* z, hipStream_t st, const int* d_done) = 0;
};

}  // namespace hg
