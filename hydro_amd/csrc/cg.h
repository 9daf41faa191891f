// cg.h -- conjugate-gradient pressure solvers on the GPU (host interface).
//
// The operator is the SIMPLE pressure-correction matrix (fluid.hpp:946-1027)
// in face form: (A p)_c = sum_f k_f (p_c - p_nb), k_f > 0 on inner faces,
// k_f = 0 on walls, so A is symmetric positive semi-definite (singular, with
// the constants as null space, when no pressure point is fixed).  Given as
// kx[(nx+1)*ny], ky[nx*(ny+1)].  Right-hand sides are projected to zero mean
// when the system is singular.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace hg {

class Precond;

struct CgResult {
  int iterations = 0;     // total inner iterations
  int outer = 0;          // refinement rounds (mixed)
  double rel_residual = 0.;  // true FP64 ||b - A x|| / ||b||
  bool converged = false;
  double ms = 0.;
  bool used_fallback = false;  // learned preconditioner rejected
  double tol = 0.;             // the solve's tolerance (for a deferred check)
  bool residual_pending = false;  // rel_residual and converged not read back yet
};

class CgSolver {
 public:
  CgSolver(int nx, int ny, hipStream_t stream);
  ~CgSolver();
  CgSolver(const CgSolver&) = delete;
  CgSolver& operator=(const CgSolver&) = delete;

  // Face coefficients already in device memory (FP64), multiplied by `sign`
  // (+1 or -1: exact) on the way in.  One pass writes the FP64 operator, its
  // FP32 copy and the FP64 and FP32 inverse diagonals.  The inputs must not
  // alias this solver's own arrays.
  void SetOperator(const double* d_kx, const double* d_ky, double sign = 1.);
  // The FP32 face coefficients and inverse diagonal, rewritten by every
  // SetOperator; the multigrid's finest level reads them in place.
  float* kx32() const { return kx32_; }
  float* ky32() const { return ky32_; }
  float* dinv32() const { return dinv32_; }
  // The FP64 face coefficients and inverse diagonal (the FP64 V-cycle's
  // finest level reads them in place).
  double* kx64() const { return kx_; }
  double* ky64() const { return ky_; }
  double* dinv64() const { return dinv_; }
  // Solves A x = b (device pointers, FP64) from x = 0, or, with warm, from
  // the x the caller left in d_x (flexible and FP64 CG; the mixed
  // refinement always starts from 0).  Convergence is relative to ||b||.
  // precision: 0 FP64 Jacobi-PCG; 1 FP32 inner PCG under FP64 refinement.
  // With `pc` set: flexible CG with FP64 Krylov vectors preconditioned by
  // pc (FP32 output); falls back to Jacobi-PCG when pc breaks descent.
  // With defer_final the closing true-residual check and the solve timer are
  // queued without a wait: `iterations` is final at return, rel_residual,
  // converged and ms come from FinishDeferred once the caller has
  // synchronised the stream (before the next Solve).
  CgResult Solve(const double* d_b, double* d_x, double tol, int max_iters, int precision,
                 bool project, Precond* pc = nullptr, bool defer_final = false,
                 bool warm = false);
  void FinishDeferred(CgResult& r);
  // y = A x (FP64), for tests and residual checks.
  void Apply(const double* d_x, double* d_y);
  const double* kx() const { return kx_; }
  const double* ky() const { return ky_; }
  int nx() const { return nx_; }
  int ny() const { return ny_; }
  hipStream_t stream() const { return stream_; }

 private:
  int nx_, ny_;
  long n_;
  hipStream_t stream_;
  double *kx_ = nullptr, *ky_ = nullptr, *dinv_ = nullptr;
  float *kx32_ = nullptr, *ky32_ = nullptr, *dinv32_ = nullptr;
  // FP64 work vectors
  double *r_ = nullptr, *z_ = nullptr, *pa_ = nullptr, *pb_ = nullptr, *q_ = nullptr;
  double* b_ = nullptr;
  // FP32 work vectors
  float *r32_ = nullptr, *z32_ = nullptr, *pa32_ = nullptr, *pb32_ = nullptr, *q32_ = nullptr;
  float *e32_ = nullptr, *zprev32_ = nullptr;
  double* part_ = nullptr;   // block partials
  void* scal_ = nullptr;     // device scalars (CgScalars)
  void* host_scal_ = nullptr;  // pinned mirror
  int blocks_;
  int last_fcg_iters_ = 5;  // iterations of the last flexible solve
  double* dirx_ = nullptr;  // direction-dot partials written by the preconditioner
  int dirx_cap_ = 0;        // ... partials per dot it holds
  hipEvent_t ev0_, ev1_;

  CgResult SolveFp64(const double* b, double* x, double tol, int max_iters, bool project,
                     bool warm);
  CgResult SolveMixed(const double* b, double* x, double tol, int max_iters, bool project,
                      Precond* pc);
  CgResult SolveFlexible(const double* b, double* x, double tol, int max_iters, bool project,
                         Precond* pc, bool warm);
  void Prepare(const double* b, double* x, bool project, bool warm);
  int JacobiPcg64(double* x, int max_iters);
  int InnerFp32(double tol_inner, int max_iters, Precond* pc, bool* pc_failed);
  double ResidualFp64(const double* b, const double* x, double* r);  // returns ||r||^2
  // r = b - A x, (r, r) into the device scalars, scalars copied to the host
  // mirror: queued, no wait.
  void QueueFinalResidual(const double* b, const double* x, double* r);
  double Dot64(const double* a, const double* b);
  void ReadScalars();
};

}  // namespace hg
