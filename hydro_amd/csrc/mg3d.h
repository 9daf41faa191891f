// mg3d.h -- flexible CG with an FP32 multigrid V-cycle for the 3D step's
// 7-point pressure rows (linear_solver_pressure "fcg_mg" on hg3_*).
//
// The rows are the pressure-correction system the 3D step assembles
// (fluid.hpp:966-994, solved at fluid.hpp:1030; entries in the order back,
// south, west, centre, east, north, front): symmetric, with the
// off-diagonals of one sign and the diagonal of the other.  The
// V-cycle works on the face form k_f = -sign(diag) * offdiag (so k_f >= 0,
// walls 0) and its output is multiplied by sign(diag), an approximation of
// A^-1 for either sign convention.
//
// Levels halve all three dimensions while they are even and the level has
// more than 64 cells; the coarsest (at most kCoarsestMax cells) is solved
// approximately by Jacobi-PCG in one workgroup, and the levels below the
// finest of at most 4096 cells run with it in that one workgroup.  Coarse face coefficient =
// half the sum of the four fine faces it covers (the rediscretised operator:
// in 3D k_f = area / (distance a_f) scales with h), restriction = sum of the
// eight children's residuals, prolongation = cell-centred trilinear weighted
// by the coarse face coefficients (collapsing to the parent across walls and
// weak couplings), smoother = three Chebyshev-damped Jacobi sweeps on
// [1/5, 2] before and after the coarse correction.
//
// Flexible CG (Notay's FCG(1)) runs in FP64 on the rows; every kernel of an
// iteration tests the stop flag, so iterations queued past convergence
// return at once.  The host queues iterations in batches and reads the
// scalars between batches.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace hg3 {

struct Fcg3Result {
  int iterations = 0;
  bool converged = false;
  bool failed = false;  // (r, z) (p, A p) <= 0: the preconditioner broke descent
};

class Mg3Fcg {
 public:
  static constexpr int kCoarsestMax = 4096;
  // Whether an nx x ny x nz grid halves to a coarsest level of at most
  // kCoarsestMax cells.
  static bool Supports(int nx, int ny, int nz);
  // Allocates the hierarchy and the Krylov vectors (throws on HIP errors or
  // an unsupported grid).
  Mg3Fcg(int nx, int ny, int nz);
  ~Mg3Fcg();
  Mg3Fcg(const Mg3Fcg&) = delete;
  Mg3Fcg& operator=(const Mg3Fcg&) = delete;
  // Solves A x = b, A the rows t[0..6] (FP64, cell-major), b in the range of
  // A (mean zero for the singular all-wall operator), from x = 0 to
  // ||b - A x|| <= tol ||b|| or max_iters iterations.
  Fcg3Result Solve(double* const (&t)[7], const double* b, double* x, double tol, int max_iters,
                   hipStream_t st);
  int levels() const { return int(lv_.size()); }

 private:
  struct Level {
    int nx, ny, nz;
    long n;
    float *kx, *ky, *kz, *dinv, *r, *xa, *xb;
  };
  std::vector<Level> lv_;
  int first_fused_ = -1;  // first level (>= 1) run with the coarser ones in one workgroup
  std::vector<void*> allocs_;
  float* sgn_ = nullptr;  // sign of the finest rows' diagonal
  float* z32_ = nullptr;
  double *r_ = nullptr, *pa_ = nullptr, *pb_ = nullptr, *q_ = nullptr, *part_ = nullptr;
  void* scal_ = nullptr;       // device Fcg3 scalars
  void* host_scal_ = nullptr;  // pinned copy
  int last_iters_ = 8;
  void* Alloc(size_t bytes);
  void Build(double* const (&t)[7], hipStream_t st);
  void Cycle(int l, float* out, hipStream_t st, const int* done);
};

}  // namespace hg3
