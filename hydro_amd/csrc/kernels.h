// kernels.h -- launchers of the HIP kernels (step, solvers, learned net).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "device_common.h"

namespace hg {

// The message hg_last_error(NULL) returns on this thread (context.hip).
void SetThreadError(const std::string& msg);


// Velocity boundary values per side (wall velocity), per component.
struct Bc {
  double vel[2][4];
  // per boundary face (left j, right ny + j, bottom 2 ny + i, top
  // 2 ny + nx + i) when the values vary along a side (outlets); nullptr:
  // the per-side values above
  const double* fvel[2] = {nullptr, nullptr};
};

struct Phases {
  int n;
  double rho[3];
  double mu[3];
  double grav[2];
  double force[2];
};

// T: storage type of the gradients, the per-step face fields and the rows
// (double: FP64 path; float: mixed path).  The state arrays are FP64.
template <class T>
struct MomentumArgsT {
  const double *u0, *u1;
  const T *g0x, *g0y, *g1x, *g1y, *gpx, *gpy, *frx, *fry, *ffmu;
  const double *flux, *rho, *utp0, *utp1, *utc0, *utc1;
  double bdf0, bdf1, bdf2, relax;
  T *s, *w, *c, *e, *n, *k0, *k1;
  double* ap = nullptr;  // when set, also a_P of the relaxed rows (fluid.hpp:876-883)
};
using MomentumArgs = MomentumArgsT<double>;
using MomentumArgsF = MomentumArgsT<float>;

// Boundary conditions of the temperature, per boundary face (left j,
// right ny + j, bottom 2 ny + i, top 2 ny + nx + i): kind 0 value fixed,
// kind 1 normal derivative fixed, to val.
struct HeatFaces {
  const int* kind;
  const double* val;
};
struct HeatIn {
  const double *t;          // previous iterate
  const double *ttp, *ttc;  // time layers
  const double* flux;       // volume flux
  const double *cond, *gx, *gy;  // filled by LaunchHeatRows
  double bdf0, bdf1, bdf2, relax;
};

// ---- step kernels (step_kernels.hip) ----
// Each step kernel with intermediates has an FP64 form and a mixed form
// (float intermediates; FP64 state and residuals).
void LaunchForceFaces(hipStream_t st, const Geo& g, const double* fx, const double* fy,
                      const double* mu, double* ffex, double* ffey, double* ffmu);
void LaunchForceFaces(hipStream_t st, const Geo& g, const double* fx, const double* fy,
                      const double* mu, float* ffex, float* ffey, float* ffmu);
void LaunchRestoredForce(hipStream_t st, const Geo& g, const double* ffex, const double* ffey,
                         double* frx, double* fry);
void LaunchRestoredForce(hipStream_t st, const Geo& g, const float* ffex, const float* ffey,
                         float* frx, float* fry);
void LaunchFacesDeriv0Pair(hipStream_t st, const Geo& g, const double* ax, const double* ay,
                           double* fx, double* fy);
void LaunchFacesDeriv0Pair(hipStream_t st, const Geo& g, const float* ax, const float* ay,
                           float* fx, float* fy);
void LaunchCellGrads(hipStream_t st, const Geo& g, const Bc& bc, const double* p,
                     const double* u0, const double* u1, double* gpx, double* gpy, double* g0x,
                     double* g0y, double* g1x, double* g1y);
void LaunchCellGrads(hipStream_t st, const Geo& g, const Bc& bc, const double* p,
                     const double* u0, const double* u1, float* gpx, float* gpy, float* g0x,
                     float* g0y, float* g1x, float* g1y);
void LaunchMomentum(hipStream_t st, const Geo& g, const Bc& bc, const MomentumArgs& a);
void LaunchMomentum(hipStream_t st, const Geo& g, const Bc& bc, const MomentumArgsF& a);
void LaunchRhieChow(hipStream_t st, const Geo& g, const Bc& bc, const double* u0,
                    const double* u1, const double* gpx, const double* gpy, const double* ffrx,
                    const double* ffry, const double* ffex, const double* ffey, const double* p,
                    const double* ap, double rc, double mvx, double mvy, double* fstar,
                    double* fcoef);
void LaunchRhieChow(hipStream_t st, const Geo& g, const Bc& bc, const double* u0,
                    const double* u1, const float* gpx, const float* gpy, const float* ffrx,
                    const float* ffry, const float* ffex, const float* ffey, const double* p,
                    const double* ap, double rc, double mvx, double mvy, double* fstar,
                    double* fcoef);
void LaunchPressureRows(hipStream_t st, const Geo& g, const double* fstar, const double* fcoef,
                        int pinned, double pval, double* s, double* w, double* c, double* e,
                        double* n, double* k);
// k of the pressure rows only (no pinned cell); -k with `negate`.
void LaunchPressureRhs(hipStream_t st, const Geo& g, const double* fstar, double* k,
                       bool negate = false);
void LaunchCorrectCells(hipStream_t st, const Geo& g, const double* pc, const double* ap,
                        const double* pprev, double alpha_p, double* pcurr, double* u0,
                        double* u1, const double* up0, const double* up1, double* part,
                        double* rs);
// flux32 (may be NULL): an FP32 copy of the corrected flux (mixed path).
void LaunchCorrectFaces(hipStream_t st, const Geo& g, const double* fstar, const double* fcoef,
                        const double* pc, double* flux, float* flux32 = nullptr);
// b = float(a), n values.
void LaunchToFloat(hipStream_t st, long n, const double* a, float* b);
void LaunchInitialFlux(hipStream_t st, const Geo& g, const Bc& bc, const double* u0,
                       const double* u1, double mvx, double mvy, double* flux);
// One advection sub-step of nf (1..3) fields with gradients formed in place;
// bval (may be NULL): per field, the inlet value of each boundary face, on
// the sides with inlet[s] != 0.
void LaunchAdvStep(hipStream_t st, const Geo& g, int nf, const double* const u[3],
                   const double* flux, double dt, double* const un[3],
                   const double* const bval[3] = nullptr, const int* inlet = nullptr);
// The mixed form: FP32 flux, gradients and Superbee face values; FP64
// fields and FP64 sums of the face terms.
void LaunchAdvStep(hipStream_t st, const Geo& g, int nf, const double* const u[3],
                   const float* flux, double dt, double* const un[3],
                   const double* const bval[3] = nullptr, const int* inlet = nullptr);
// Inlets and outlets (kind[s]: hg_bc_kind per side).  Inlet values of the
// phases into pdb[p * nb + b] (nb = 2 (nx + ny) boundary faces); the outlet
// update of the boundary velocities bv0 / bv1 from the cells' u0 / u1
// (FluidSimple::UpdateOutletBaseConditions); scratch holds 2 nb + 1 values.
void LaunchGatherInlet(hipStream_t st, const Geo& g, const int kind[4], const double* const pd[3],
                       int np, double* pdb);
void LaunchOutletUpdate(hipStream_t st, const Geo& g, const int kind[4], const double* u0,
                        const double* u1, double* bv0, double* bv1, double* scratch);
void LaunchProperties(hipStream_t st, const Geo& g, const Phases& ph, const double* const pd[3],
                      double* const alpha[3], double* rho, double* mu, double* fx, double* fy);
void LaunchSmooth(hipStream_t st, const Geo& g, const double* u, double* out);
// Two smoothing passes in one launch (out must not alias u).
void LaunchSmooth2(hipStream_t st, const Geo& g, const double* u, double* out);
void LaunchPhase0Init(hipStream_t st, const Geo& g, const Phases& ph, const double* pd1,
                      const double* pd2, double* pd0);
void LaunchNanCheck(hipStream_t st, const double* a, const double* b, const double* c, int n,
                    int* flag);
// The temperature's rows of one heat iteration (heat.hpp:66-75,
// conv_diff.hpp:130-243) in delta form: cell conductivity (from the volume
// fractions and conductivity_i) and the temperature gradient into cond, gx,
// gy, then the 5-point rows into out / k.
void LaunchHeatRows(hipStream_t st, const Geo& g, const HeatFaces& hb, const double* const alpha[3],
                    int np, const double cond_i[3], const HeatIn& in, double* cond, double* gx,
                    double* gy, double* s, double* w, double* c, double* e, double* n,
                    double* k);
// CalcStat's boundary in/out integrals: acc[4p..4p+3] (volume in, out, mass
// in, out of phase p) += dt * the boundary fluxes of the time_curr flux;
// part holds 6 * BoundaryFluxBlocks(g) partial sums.
int BoundaryFluxBlocks(const Geo& g);
void LaunchBoundaryFlux(hipStream_t st, const Geo& g, const double* flux,
                        const double* const alpha[3], int np, const double density[3], double dt,
                        double* part, double* acc);
// FluidSimple::GetAutoTimeStep into out[0]; part holds StatsBlocks(g) values.
void LaunchAutoDt(hipStream_t st, const Geo& g, const double* flux, double* part, double* out);
int StatsBlocks(const Geo& g);
void LaunchStats(hipStream_t st, const Geo& g, const double* xn, const double* yn,
                 const double* alpha, const double* pd, const double* u0, const double* u1,
                 double* part, double* out7);

// ---- 5-point solvers (solvers.hip) ----
// Rows: s x[r-nx] + w x[r-1] + c x[r] + e x[r+1] + n x[r+nx] + k = 0.
template <class T>
struct RowsT {
  const T *s, *w, *c, *e, *n;
};
using Rows = RowsT<double>;
using RowsF = RowsT<float>;

// Exact-order symmetric Gauss-Seidel sweep from zero (linear.hpp:527-567) for
// up to two right-hand sides sharing the coefficients, in one workgroup.
// x_out[r] = base[r] + x[r] when base != nullptr (the delta-form update).
void LaunchLuWavefront(hipStream_t st, int nx, int ny, Rows a, const double* k0,
                       const double* k1, double* x0, double* x1, const double* base0,
                       const double* base1, double* out0, double* out1);
// Exact-order lexicographic SOR from zero (linear.hpp:685-715), all sweeps in
// one persistent workgroup.  info[0] = iterations (reference count),
// info[1] = final max|corr| (as double bits).
void LaunchGsWavefront(hipStream_t st, int nx, int ny, Rows a, const double* k, double* x,
                       double tol, int limit, double omega, unsigned long long* info);
// LuDecompositionRelaxed (linear.hpp:578-651) from zero, every iteration in
// one persistent workgroup: res receives the solution; corr and f are
// scratch cell arrays; info as for LaunchGsWavefront.
void LaunchLuRelaxed(hipStream_t st, int nx, int ny, Rows a, const double* k, double* res,
                     double* corr, double* f, double tol, int limit, double relax,
                     unsigned long long* info);
// Jacobi (linear.hpp:750-783), batched sweeps with a device-side stop test.
struct JacobiWork {
  double* xa;
  double* xb;
  unsigned long long* diff;  // [max_sweeps + 1]
  int* state;                // [0] done flag, [1] counter, [2] iterations, [3] result buffer
};
void LaunchJacobiSweeps(hipStream_t st, int nx, int ny, Rows a, const double* k, JacobiWork w,
                        int first_sweep, int nsweeps, double tol, int limit, double omega);
// Red-black ordered symmetric GS sweep ("lu_rb"), two right-hand sides,
// out = base + x; the float form solves the correction in FP32 (mixed path).
void LaunchLuRb(hipStream_t st, int nx, int ny, Rows a, const double* k0, const double* k1,
                double* xb0, double* xb1, const double* base0, const double* base1,
                double* out0, double* out1);
void LaunchLuRb(hipStream_t st, int nx, int ny, RowsF a, const float* k0, const float* k1,
                float* xb0, float* xb1, const double* base0, const double* base1,
                double* out0, double* out1);

// SIMPLER's second pressure right-hand side (fluid.hpp:1059-1151) into k,
// from the velocity rows and constants of the iteration, the corrected and
// previous velocity, restored force, pressure gradient, face force, a_P, the
// corrected flux and the pressure rows at the current pressure p; ev0 / ev1
// are scratch cell arrays.
void LaunchSimpler(hipStream_t st, const Geo& g, Rows vel, const double* k0, const double* k1,
                   const double* uc0, const double* uc1, const double* up0, const double* up1,
                   const double* frx, const double* fry, const double* gpx, const double* gpy,
                   const double* ffex, const double* ffey, const double* ap, const double* flux,
                   double rc, Rows pres, const double* p, int pinned, double pval, double* ev0,
                   double* ev1, double* k);

// ---- Poisson / CG (cg.hip) ----
struct CgSystem;  // opaque, defined in cg.hip

}  // namespace hg
