/*
 * hydro_noalloc.h -- the allocation-free variant of hg_stencil_solve
 * (libhydrogpu.so).
 *
 * No counterpart in the reference.  It is for callers that cannot allocate
 * device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  It runs the same kernels as
 * hg_stencil_solve, whose rows replace LinearSolver::Solve
 * (linear.hpp:302-310), on arrays the caller already holds on the device.
 */
#ifndef HYDRO_NOALLOC_H_
#define HYDRO_NOALLOC_H_

#include "hydro_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Size of the library's file-scope device work array.  Each GPU the
 * library runs on holds one until the process ends. */
#define HG_NOALLOC_SCRATCH_BYTES (16u << 20)

/* hg_stencil_solve on device arrays: s, w, c, e, n, k and x are device
 * pointers of nx*ny doubles on the current device (same row meaning, same
 * zero initial guess, same result and iteration count as hg_stencil_solve).
 * solver: lu, lu_rb, gauss_seidel, jacobi or lu_relaxed.  The work arrays
 * come from the library's scratch array; no hipMalloc, hipFree or
 * allocating library routine is called.  Needs
 *   256 + 8 * (max_iters + 2) + 4 * 8 * nx * ny   (each part rounded up to
 *   256 bytes) <= HG_NOALLOC_SCRATCH_BYTES
 * and nx <= 4096, else HG_ERR_BAD_ARG (hg_last_error(NULL) says why).
 * stream: the hipStream_t to run on (NULL: the null stream); the call waits
 * for it before it returns.  Every call on one GPU shares that GPU's scratch
 * array and leaves the last call's contents there: calls that overlap on one
 * GPU race.  stats: iterations, final_diff and converged as in
 * hg_stencil_solve; rel_residual is -1 (not computed) and ms 0. */
hg_status hg_stencil_solve_noalloc(int32_t nx, int32_t ny, const double* s, const double* w,
                                   const double* c, const double* e, const double* n,
                                   const double* k, double* x, const hg_solve_opts* opts,
                                   hg_solve_stats* stats, void* stream);

#ifdef __cplusplus
}
#endif

#endif  /* HYDRO_NOALLOC_H_ */
