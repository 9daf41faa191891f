// learned.h -- preconditioner interface of the flexible CG solver.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.h"

namespace hg {

// Flexible-CG work a preconditioner can take over in the first and the last
// kernel of its application (Precond::ApplyFcg), saving two launches per
// iteration.
struct FcgFusion {
  // Stop test of iteration k, by workgroup 0 of the first kernel: rr = sum
  // of the nb partials of (r, r); *iter = k; *done = 1 when rr <= tol2 * bb,
  // k >= max_iter, rr is NaN or *fail.  Other workgroups of that kernel carry
  // on (their work is discarded); every later kernel sees *done.
  const double* rr_part;
  int nb, k;
  double* rr;
  int *iter, *done;
  const double *tol2, *bb;
  const int *max_iter, *fail;
  // Direction dots, by the last kernel: per-workgroup partials of
  // (z, q_old) at dir[b] and (r, z) at dir[G + b], G its grid size, both
  // with z rounded to FP32 as written.  q_old is null at k = 0 ((z, q_old)
  // is then 0).
  const double* qold;
  double* dir;
  // The previous iteration's update, folded into the finest down-leg
  // (up_on): with alpha = *up_alpha, x += alpha p and r -= alpha q on the
  // kernel's tile cells (the new r written to up_r, not over its input),
  // and r - alpha q on its halo as the down-leg's input; (r, r) of the tile
  // cells at up_rr[b].  The stop test then runs in the next level's first
  // kernel.
  int up_on;
  const double* up_alpha;
  const double *up_p, *up_q;
  double *up_x, *up_r, *up_rr;
};

// The stop test above, by one workgroup of 256 threads.
__device__ __forceinline__ void FcgStopTest(const FcgFusion& f) {
  __shared__ double red[4];
  double s = 0.;
  for (int b = threadIdx.x; b < f.nb; b += 256) s += f.rr_part[b];
  s = WaveSum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double rr = ((red[0] + red[1]) + red[2]) + red[3];
    *f.rr = rr;
    *f.iter = f.k;
    if (!(rr > *f.tol2 * *f.bb) || f.k >= *f.max_iter || !(rr == rr) || *f.fail) *f.done = 1;
  }
}

// z = M(r): r is the FP64 residual on the cell grid, z the FP32 correction.
// Kernels must return at once when *d_done is non-zero (solve converged).
class Precond {
 public:
  virtual ~Precond() {}
  virtual void Apply(const double* r, float* z, hipStream_t st, const int* d_done) = 0;
  // Apply plus the FcgFusion work; *ndir = the number of partials written
  // per dot.  Returns false, doing nothing, where not supported.
  virtual bool ApplyFcg(const double* r, float* z, hipStream_t st, const int* d_done,
                        const FcgFusion& f, int* ndir) {
    (void)r, (void)z, (void)st, (void)d_done, (void)f, (void)ndir;
    return false;
  }
  // Partials per dot that ApplyFcg writes (0: not supported).
  virtual int FcgPartials() const { return 0; }
  // ApplyFcg takes the previous iteration's update (FcgFusion::up_on).
  virtual bool FusesUpdate() const { return false; }
};

}  // namespace hg
