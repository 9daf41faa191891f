// mg.h -- FP32 cell-centred multigrid V-cycle for the pressure operator.
//
// Face-form operator (A p)_c = sum_f k_f (p_c - p_nb).  Coarse levels halve
// both dimensions while they stay even: coarse face coefficient = mean of
// the two fine faces it covers (rediscretisation; in 2D k_f = A/(d a_f) does
// not scale with h), restriction = sum of the four children's residuals,
// prolongation = cell-centred bilinear weighted by the coarse face
// coefficients (injection across walls and strong density jumps), smoother =
// Chebyshev-damped Jacobi sweeps before and after the coarse correction (two
// on the finest level, three on the coarser ones),
// levels of <= 2048 cells in one workgroup down to a coarsest grid of <= 64
// cells solved by Jacobi-PCG in one wave.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "learned.h"
#include "tile.h"

namespace hg {

constexpr int kTallY = 16;  // tile height of the finest level's fused kernels
constexpr int kMidY = 4;    // tile height of the coarser levels' fused kernels

// A level's arrays, in the level type M (float: the V-cycle of the mixed
// path; double: the same cycle in FP64, for comparison).
template <class M>
struct MgLevelT {
  int nx, ny;
  TileGrid tg;    // 64 x 4 tiles
  TileGrid mid;   // 64 x kMidY tiles
  TileGrid tall;  // 64 x kTallY tiles
  M *kx = nullptr, *ky = nullptr, *dinv = nullptr;
  M *r = nullptr, *xa = nullptr, *xb = nullptr;
};
using MgLevel = MgLevelT<float>;

// Damped Jacobi sweeps of each smoothing pass (pre and post): kFineSweeps on
// the finest level, kSweeps on the coarser ones; a fused kernel's tile
// carries a halo of as many cells as it sweeps.
constexpr int kFineSweeps = 2;
constexpr int kSweeps = 3;
constexpr int kMaxSweeps = 3;
struct SweepWeights {
  float w[kMaxSweeps];  // damping of sweep 1, 2, ... (those past the count unused)
};

// The coarse levels handed to the one-workgroup cycle kernel.

constexpr int kMaxInKernelLevels = 12;
constexpr int kLdsCells = 2752;  // LDS cells per vector for all in-kernel levels
template <class M>
struct CoarseSetT {
  int count, bilinear;
  int off[kMaxInKernelLevels];  // offset of each level in the LDS vectors
  int nx[kMaxInKernelLevels], ny[kMaxInKernelLevels];
  const M *kx[kMaxInKernelLevels], *ky[kMaxInKernelLevels], *dinv[kMaxInKernelLevels];
  M *r[kMaxInKernelLevels], *xa[kMaxInKernelLevels], *xb[kMaxInKernelLevels];
  M* out;  // the first level's correction (the caller's `out`)
};

// The V-cycle with its levels' arrays and arithmetic in M (the output z is
// FP32 either way: the Precond interface's).  MgPrecond (M = float) is the
// preconditioner of fcg_mg / fcg_learned; MgPrecond64 runs the same kernels
// in FP64 to measure what FP32 buys.
template <class M>
class MgPrecondT : public Precond {
 public:
  // kx0, ky0, dinv0: the finest level's face coefficients and inverse
  // diagonal in M, owned by the caller (CgSolver::kx32() ... for float,
  // kx64() ... for double) and rewritten by it before each Build.
  MgPrecondT(int nx, int ny, M* kx0, M* ky0, M* dinv0);
  ~MgPrecondT();
  MgPrecondT(const MgPrecondT&) = delete;
  MgPrecondT& operator=(const MgPrecondT&) = delete;
  // Coarse-level coefficients and inverse diagonals from the finest level's.
  void Build(hipStream_t st);
  // z = V(r): r is the FP64 residual of the finest level; z FP32.  Does
  // nothing when *d_done is set.
  void Apply(const double* r, float* z, hipStream_t st, const int* d_done) override;
  // Apply with flexible CG's stop test in the finest level's first kernel
  // and its direction dots in the finest level's last kernel.
  bool ApplyFcg(const double* r, float* z, hipStream_t st, const int* d_done,
                const FcgFusion& f, int* ndir) override;
  int FcgPartials() const override { return lv_.size() > 1 ? lv_[0].tall.blocks() : 0; }
  // level 1 runs a tiled down-leg (its first kernel takes the stop test)
  bool FusesUpdate() const override {
    return lv_.size() >= 3 && !(use_coarse_kernel && first_in_kernel_ == 1);
  }
  int levels() const { return int(lv_.size()); }
  // Average device time (ms) of one launch of a V-cycle kernel ("KDown<double>",
  // "KUp<double, true>" on the finest level, "KCoarseCycle"), launched `reps`
  // times back to back between two HIP events on `st`, reading r (FP64,
  // finest level) and the level buffers.  Throws on an unknown name.
  double TimeKernel(const char* name, int reps, const double* r, hipStream_t st);
  const MgLevelT<M>& level(int l) const { return lv_[l]; }
  float omega = 0.8f;  // damping of the plain Jacobi sweeps (coarsest-level fallbacks)
  // Chebyshev smoothing (pre and post): damping 1/r_k at the roots r_k of
  // the Chebyshev polynomial of the sweep count on an interval [a, 2] of the
  // spectrum of D^-1 A (which is in [0, 2] by Gershgorin), applied in that
  // order (small damping first).
  // Finest level, two sweeps on [1/2, 2]: r_k = 5/4 + 3/4 cos((2k-1) pi / 4),
  // the sweeps' product at most 1/T_2(5/3) = 0.22 on that interval.
  SweepWeights fine_sweeps{{0.561695f, 1.389532f, 0.f}};
  // Coarser levels, three sweeps on [1/5, 2]: r_k = 11/10 + 9/10
  // cos((2k-1) pi / 6), the product at most 1/T_3(11/9) = 0.27 on an
  // interval that two sweeps ([1/2, 2]) leave undamped below 1/2.  These
  // levels are latency-bound, so the third sweep (in LDS, one ring more of
  // halo) costs little; on the Rayleigh-Taylor pressure systems it takes
  // flexible CG from 7.35 to about 6.2 iterations per solve to 1e-8, where
  // a third sweep on the finest level too adds nothing more.
  SweepWeights sweeps{{0.532078f, 0.909091f, 3.119374f}};
  int coarse_sweeps = 40;
  bool bilinear = true;  // prolongation: bilinear (true) or injection
  // levels of <= 2048 cells down to a coarsest grid of <= 64 cells run in
  // one workgroup (one launch) instead of one launch per sweep
  bool use_coarse_kernel = true;

 private:
  std::vector<MgLevelT<M>> lv_;
  int first_in_kernel_ = -1;
  // a device flag that stays 0: the stop flag of cycles run without one, so
  // that the kernels always read a flag (no null test ahead of their loads)
  int* never_done_ = nullptr;
  // level l's correction: into z (FP32) on level 0, into out on the others
  void Cycle(int l, const double* r64, float* z, M* out, hipStream_t st, const int* d_done,
             const FcgFusion* f = nullptr);
};
using MgPrecond = MgPrecondT<float>;
using MgPrecond64 = MgPrecondT<double>;

}  // namespace hg
