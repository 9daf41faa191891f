// context.hip -- the C-ABI of libhydrogpu.so: context lifecycle, field
// transfer and the time step driven at the reference's granularity
// (hydro<Mesh>::step, hydro2d.hpp:1531-1621; FluidSimple, fluid.hpp:421-1207;
// AdvectionSolverMultiExplicit, advection.hpp:368-555).
//
// All fields live in HBM for the life of the context; the host only
// sequences kernels on the context's stream and reads back the few scalars
// the reference's control flow branches on (the SIMPLE convergence
// indicator, solver iteration counts, NaN flags).

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#include <thread>
#include <atomic>
#include <chrono>

#include "../../include/hydro_gpu.h"
#include "cg.h"
#include "host_geometry.h"
#include "kernels.h"
#include "learned_net.h"
#include "mg.h"

#define HG_CHECK(x)                                                                  \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") +            \
                                                   hipGetErrorString(e_) + " at " #x); \
  } while (0)

namespace {

struct HgError : std::runtime_error {
  hg_status code;
  HgError(hg_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// GetDerivativeApproxCoeffs (solver.hpp:815-861): Lagrange weights of the
// time derivative at 0 from samples at args (BDF2 / BDF1).
std::vector<double> DerivCoeffs(double target, const std::vector<double>& args) {
  size_t size = args.size();
  std::vector<double> res(size);
  for (size_t i = 0; i < size; ++i) {
    double denom = 1., numer = 0.;
    for (size_t j = 0; j < size; ++j) {
      if (j == i) continue;
      denom *= args[i] - args[j];
      double term = 1.;
      for (size_t k = 0; k < size; ++k)
        if (k != i && k != j) term *= target - args[k];
      numer += term;
    }
    res[i] = numer / denom;
  }
  return res;
}

// ---- small elementwise kernels -------------------------------------------
// Initial guess of a step (fluid.hpp:800-811, conv_diff.hpp:122-128):
// iter_curr = time_curr + (time_curr - time_prev) * guess_extrapolation, for
// u0, u1, p (threads < nc) and the volume flux (the next nf threads) in one
// launch.
struct ExtrapLayers {
  const double *tc[4], *tp[4];
  double* out[4];
};

__global__ void KExtrapolate(long nc, long nf, ExtrapLayers l, double ge) {
  const long i = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (ge == 0.) {
    // a + (a - b) * 0 is a itself, except for a = -0 (the sign of the sum
    // follows b) and a non-finite: only there is time_prev read
    auto Keep = [](double a, const double* tp, long k) {
      return (a == 0. || !isfinite(a)) ? a + (a - tp[k]) * 0. : a;
    };
    if (i < nc) {
      double a[3];
      for (int k = 0; k < 3; ++k) a[k] = l.tc[k][i];
      for (int k = 0; k < 3; ++k) l.out[k][i] = Keep(a[k], l.tp[k], i);
    } else if (i < nc + nf) {
      const long f = i - nc;
      l.out[3][f] = Keep(l.tc[3][f], l.tp[3], f);
    }
    return;
  }
  if (i < nc) {
    double a[3], b[3];
    for (int k = 0; k < 3; ++k) {
      a[k] = l.tc[k][i];
      b[k] = l.tp[k][i];
    }
    for (int k = 0; k < 3; ++k) l.out[k][i] = a[k] + (a[k] - b[k]) * ge;
  } else if (i < nc + nf) {
    const long f = i - nc;
    const double a = l.tc[3][f], b = l.tp[3][f];
    l.out[3][f] = a + (a - b) * ge;
  }
}

// inout += x.
__global__ void KAddInPlace(long n, const double* __restrict__ x, double* __restrict__ inout) {
  const long i = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (i < n) inout[i] = inout[i] + x[i];
}

// One field's initial guess: out = tc + (tc - tp) * ge.
__global__ void KGuess(long n, const double* __restrict__ tc, const double* __restrict__ tp,
                       double ge, double* __restrict__ out) {
  const long i = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (i < n) {
    const double a = tc[i], b = tp[i];
    out[i] = a + (a - b) * ge;
  }
}

// Pressure solvers that take the operator in face form (CgSolver).
inline bool FaceSolver(int32_t s) {
  return s == HG_SOLVER_CG || s == HG_SOLVER_FCG_MG || s == HG_SOLVER_FCG_LEARNED;
}

__global__ void KNegateFaces(long nf, const double* __restrict__ fcoef, double* __restrict__ k) {
  long f = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (f < nf) k[f] = -fcoef[f];
}
__global__ void KNeg(long n, const double* __restrict__ a, double* __restrict__ b) {
  long i = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (i < n) b[i] = -a[i];
}
__global__ void KAddTo(long n, const double* __restrict__ base, const double* __restrict__ x,
                       double* __restrict__ out) {
  long i = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (i < n) out[i] = base[i] + x[i];
}

inline dim3 G(long n) { return dim3(unsigned((n + 255) / 256)); }

__global__ void KNegateF(long n, float* __restrict__ z) {
  const long i = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (i < n) z[i] = -z[i];
}

// The check knob fcg_negate_precond_gpu: minus the wrapped preconditioner,
// applied without the fused FCG kernels, so that FCG's descent check fails
// on its first iteration and the solve takes the Jacobi-PCG fallback.
class NegatedPrecond : public hg::Precond {
 public:
  NegatedPrecond(hg::Precond* inner, long n) : inner_(inner), n_(n) {}
  void Apply(const double* r, float* z, hipStream_t st, const int* d_done) override {
    inner_->Apply(r, z, st, d_done);
    KNegateF<<<G(n_), 256, 0, st>>>(n_, z);
  }

 private:
  hg::Precond* inner_;
  long n_;
};

}  // namespace

struct hg_ctx {
  hg_config cfg{};
  hg::HostMesh mesh;
  hg::Geo geo{};
  hg::Bc bc{};
  hg::Phases ph{};
  hipStream_t st = nullptr;
  long nc = 0, nf = 0;
  int np = 1;
  std::vector<void*> allocs;
  // layers
  double *u_tc[2], *u_tp[2], *u_ic[2], *u_ip[2];
  double *p_tc, *p_tp, *p_ic, *p_ip;
  double *fl_tc, *fl_tp, *fl_ic, *fl_ip;
  // buffer freed by the last FinishStep's rotation (time layers)
  double *u_spare[2] = {nullptr, nullptr}, *p_spare = nullptr, *fl_spare = nullptr;
  // the four buffers of each layered field (u0, u1, p, flux)
  double* layer_bufs[4][4] = {};
  // The mixed path with guess_extrapolation 0 takes the step's guess in
  // place: the first SIMPLE iteration reads time_curr itself as its previous
  // iterate instead of a copy of it (the copy is KExtrapolate's whole work
  // when the extrapolation is 0)
  bool guess_in_place = false;
  // properties
  double *rho, *rho_s, *mu, *mu_s, *fx, *fy, *tmp1, *tmp2;
  double *pd[3], *pd_new[3], *alpha[3];
  // per-step faces
  double *ffex, *ffey, *ffmu, *frx, *fry, *ffrx, *ffry;
  // per-iteration
  double *gpx, *gpy, *g0x, *g0y, *g1x, *g1y;
  double *As, *Aw, *Ac, *Ae, *An, *k0, *k1, *x0, *x1;
  double *ap, *fstar, *fcoef;
  double *Ps, *Pw, *Pc, *Pe, *Pn, *Pk, *pcorr, *pb, *kface, *zeros;
  double *xn_d, *yn_d, *stat_part, *stat_out, *rs_part;
  unsigned long long* rs_d;
  int* nan_d;
  unsigned long long* gs_info;
  hg::JacobiWork jw{};
  double* jdiff_host_dummy = nullptr;
  // pinned host mirrors
  unsigned long long* rs_h = nullptr;
  int* flag_h = nullptr;
  unsigned long long* info_h = nullptr;
  double* stat_h = nullptr;
  int pinned = -1;
  double bdf[3] = {0, 0, 0};
  double t_fluid = 0., t_adv = 0., dt = 0., dt_adv = 0.;
  // the module's own clock (TModule::tn / tnp, module.cpp:74-81), advanced
  // once per step with the dt in force after it
  double mod_t = 0., mod_tnp = 0.;
  int fluid_iter = 0, cd_iter = 0;
  double last_rs = 1.;
  hg_iter_stats last_iter{};
  // the module state the output reads: "s" (the last iteration count), the
  // steps finished, the Rs of the current step's iterations, and the running
  // part of CalcStat (hydro2d.hpp:1474-1528): boundary in/out integrals per
  // phase on the device (volume_in, volume_out, mass_in, mass_out), the
  // output offset meshpos and the value of it the last statistics used
  int last_s = 0;
  long steps = 0;
  // heat (heat.hpp:18-94): temperature time layers, two scratch layers, the
  // per-boundary-face conditions, cell conductivity and temperature
  // gradient, the unsteady coefficients at the heat solver's own dt
  double *t_tc = nullptr, *t_tp = nullptr, *t_a = nullptr, *t_b = nullptr;
  int* hb_kind = nullptr;
  double* hb_val = nullptr;
  double *kcond = nullptr, *tgx = nullptr, *tgy = nullptr;
  double hbdf[3] = {0., 0., 0.};
  // inlets and outlets: per-boundary-face velocities (bc.fvel points into
  // bvel_d; outlets are updated every iteration), the inlet partial
  // densities, scratch of the outlet update
  bool open_bc = false, any_inlet = false, any_outlet = false;
  int side_kind[4] = {0, 0, 0, 0};
  double *bvel_d = nullptr, *pdb_d = nullptr, *obc_scratch = nullptr;
  std::vector<double> iter_rs;
  double *bflux_acc = nullptr, *bflux_part = nullptr;
  double meshpos[2] = {0., 0.}, meshpos_stat[2] = {0., 0.};
  std::string err;
  // The mixed path (precision_gpu mixed): the per-step face fields, the cell
  // gradients, the momentum rows and the velocity correction in FP32, and an
  // FP32 copy of the corrected flux for the advection; the state, the row
  // constants' arithmetic, the pressure system and the corrections stay FP64.
  bool mixed = false;
  struct F32 {
    float *ffex, *ffey, *ffmu, *frx, *fry, *ffrx, *ffry;
    float *gpx, *gpy, *g0x, *g0y, *g1x, *g1y;
    float *s, *w, *c, *e, *n, *k0, *k1, *xb0, *xb1;
    float* flux;
  } f32{};
  bool flux32_valid = false;  // f32.flux holds float(fl_ic)
  std::unique_ptr<hg::CgSolver> cg;
  std::unique_ptr<hg::LearnedNet> net;
  std::unique_ptr<hg::MgPrecond> mg;
  std::unique_ptr<hg::Precond> negated;  // fcg_negate_precond_gpu
  bool pressure_system = false;  // cg/mg hold the operator of a pressure solve
  hipEvent_t ev[4];
  bool timed = false;  // ev[0..3] hold the regions of the last hg_step
  hipEvent_t kev[2];   // hg_time_kernel's own pair

  double* Alloc(long n) {
    void* p = nullptr;
    HG_CHECK(hipMalloc(&p, size_t(std::max<long>(n, 1)) * sizeof(double)));
    HG_CHECK(hipMemsetAsync(p, 0, size_t(std::max<long>(n, 1)) * sizeof(double), st));
    allocs.push_back(p);
    return (double*)p;
  }
  float* AllocF(long n) {
    void* p = nullptr;
    HG_CHECK(hipMalloc(&p, size_t(std::max<long>(n, 1)) * sizeof(float)));
    HG_CHECK(hipMemsetAsync(p, 0, size_t(std::max<long>(n, 1)) * sizeof(float), st));
    allocs.push_back(p);
    return (float*)p;
  }
  ~hg_ctx() {
    if (st) (void)hipStreamSynchronize(st);
    for (void* p : allocs) (void)hipFree(p);
    if (rs_h) (void)hipHostFree(rs_h);
    if (flag_h) (void)hipHostFree(flag_h);
    if (info_h) (void)hipHostFree(info_h);
    if (stat_h) (void)hipHostFree(stat_h);
    cg.reset();
    net.reset();
    mg.reset();
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (auto& e : kev) (void)hipEventDestroy(e);
    if (st) (void)hipStreamDestroy(st);
  }
  // Readbacks that wait for the next synchronisation of the stream instead
  // of one of their own (hg_step queues them): the pressure solve's closing
  // residual check and timer, the convergence indicator.  Sync() settles
  // them into last_iter / last_rs.
  bool pending_cg = false, pending_rs = false;
  hg::CgResult cg_res;
  // p and u of time_curr were checked for NaN by FinishStep and not changed
  // since (StartStep need not check them again)
  bool fields_checked = false;
  // Warm start of the face-form pressure solve: pcorr holds the converged
  // correction of the previous step's only SIMPLE iteration, a multiple of
  // which the next step's first solve starts from (successive steps'
  // corrections are close: the pressure changes smoothly in time; at
  // 1024x2048 RT this takes FCG from 8.1 to 7.35 iterations per step, and a
  // linear extrapolation from the last two corrections gains nothing more).
  bool pcorr_converged = false, pcorr_warm = false;
  void Sync() {
    HG_CHECK(hipStreamSynchronize(st));
    Settle();
  }
  void Settle() {
    if (pending_cg) {
      pending_cg = false;
      cg->FinishDeferred(cg_res);
      last_iter.pressure_solve_diff = cg_res.rel_residual;
      last_iter.pressure_converged = cg_res.converged ? 1 : 0;
      pcorr_converged = cg_res.converged;
    }
    if (pending_rs) {
      pending_rs = false;
      double rs;
      std::memcpy(&rs, rs_h, 8);
      last_rs = rs;
      last_iter.convergence_indicator = rs;
      if (!iter_rs.empty()) iter_rs.back() = rs;  // the iteration that deferred it
    }
  }
};

namespace {

hg_status Guard(hg_ctx* ctx, const std::function<void()>& fn);

void Upload(hg_ctx* c, double* d, const std::vector<double>& h) {
  HG_CHECK(hipMemcpyAsync(d, h.data(), h.size() * 8, hipMemcpyHostToDevice, c->st));
}

void Smooth(hg_ctx* c, const double* in, double* out, int times) {
  // out = GetSmoothField(in, times); in and out may alias.
  if (times <= 0) {
    if (out != in)
      HG_CHECK(hipMemcpyAsync(out, in, c->nc * 8, hipMemcpyDeviceToDevice, c->st));
    return;
  }
  // passes two at a time in one launch, an odd last pass alone; the last
  // launch writes out directly unless out is its input
  const double* src = in;
  for (int r = 0; r < times;) {
    const int k = times - r >= 2 ? 2 : 1;
    double* dst = src == c->tmp1 ? c->tmp2 : c->tmp1;
    if (r + k == times && out != src) dst = out;
    if (k == 2)
      hg::LaunchSmooth2(c->st, c->geo, src, dst);
    else
      hg::LaunchSmooth(c->st, c->geo, src, dst);
    src = dst;
    r += k;
  }
  if (src != out)
    HG_CHECK(hipMemcpyAsync(out, src, c->nc * 8, hipMemcpyDeviceToDevice, c->st));
}

void UpdateProperties(hg_ctx* c) {
  const double* pdv[3] = {c->pd[0], c->pd[1], c->pd[2]};
  double* al[3] = {c->alpha[0], c->alpha[1], c->alpha[2]};
  hg::LaunchProperties(c->st, c->geo, c->ph, pdv, al, c->rho, c->mu, c->fx, c->fy);
  Smooth(c, c->rho, c->rho_s, c->cfg.density_smooth_times);
  Smooth(c, c->mu, c->mu_s, c->cfg.viscosity_smooth_times);
  Smooth(c, c->fx, c->fx, c->cfg.force_smooth_times);
  Smooth(c, c->fy, c->fy, c->cfg.force_smooth_times);
}

int NanCheck(hg_ctx* c, const double* a, const double* b, const double* d);
void SolveRowsIterative(hg_ctx* c, int32_t solver, hg::Rows rows, const double* k, double* x,
                        int* iters, double* diff);

// InitHeatSolver (hydro2d.hpp:652-688): a uniform temperature_initial; a
// boundary face whose centre lies in [heat_box_lb, heat_box_rt] (bounds
// included, vect.hpp:239-246) holds heat_box_temperature, every other one a
// zero normal derivative.  The heat solver keeps the dt it is built with.
void InitHeat(hg_ctx* c) {
  const hg_config& cfg = c->cfg;
  const long nc = c->nc;
  const int nx = cfg.nx, ny = cfg.ny, nb = 2 * (nx + ny);
  for (double** p : {&c->t_tc, &c->t_tp, &c->t_a, &c->t_b, &c->kcond, &c->tgx, &c->tgy})
    *p = c->Alloc(nc);
  c->hb_kind = (int*)c->Alloc(nb / 2 + 1);
  c->hb_val = c->Alloc(nb);
  std::vector<double> t0(size_t(nc), cfg.temperature_initial);
  Upload(c, c->t_tc, t0);
  Upload(c, c->t_tp, t0);
  std::vector<int> kind(static_cast<size_t>(nb));
  std::vector<double> val(static_cast<size_t>(nb));
  for (int b = 0; b < nb; ++b) {
    hg::HostMesh::Face f;
    if (b < ny) f = c->mesh.XFaceGeo(0, b);
    else if (b < 2 * ny) f = c->mesh.XFaceGeo(nx, b - ny);
    else if (b < 2 * ny + nx) f = c->mesh.YFaceGeo(b - 2 * ny, 0);
    else f = c->mesh.YFaceGeo(b - 2 * ny - nx, ny);
    const bool inside = !(f.cx < cfg.heat_box_lb[0] || cfg.heat_box_rt[0] < f.cx) &&
                        !(f.cy < cfg.heat_box_lb[1] || cfg.heat_box_rt[1] < f.cy);
    kind[size_t(b)] = inside ? 0 : 1;
    val[size_t(b)] = inside ? cfg.heat_box_temperature : 0.;
  }
  HG_CHECK(hipMemcpyAsync(c->hb_kind, kind.data(), size_t(nb) * sizeof(int),
                          hipMemcpyHostToDevice, c->st));
  Upload(c, c->hb_val, val);
  const double dt = cfg.dt;
  std::vector<double> co = DerivCoeffs(0., {-2. * dt, -dt, 0.});
  if (!cfg.time_second_order_heat) {
    std::vector<double> c2 = DerivCoeffs(0., {-dt, 0.});
    co = {0., c2[0], c2[1]};
  }
  for (int i = 0; i < 3; ++i) c->hbdf[i] = co[i];
  c->Sync();  // the host vectors above go out of scope
}

// One heat step (hydro2d.hpp:1607-1612): StartStep (guess = time_curr, the
// heat solver's extrapolation is 0), one MakeIteration (its iteration limit
// is 1), FinishStep with the "NaN field" check (conv_diff.hpp:117-128,
// 130-243, 244-251; heat.hpp:62-80).
void HeatStep(hg_ctx* c) {
  const hg_config& cfg = c->cfg;
  const int nx = c->geo.nx, ny = c->geo.ny;
  KGuess<<<G(c->nc), 256, 0, c->st>>>(c->nc, c->t_tc, c->t_tp, 0., c->t_a);
  hg::HeatFaces hb{c->hb_kind, c->hb_val};
  hg::HeatIn in{};
  in.t = c->t_a;
  in.ttp = c->t_tp;
  in.ttc = c->t_tc;
  in.flux = c->fl_ic;  // the fluid's iter_curr flux (hydro2d.hpp:685)
  in.bdf0 = c->hbdf[0];
  in.bdf1 = c->hbdf[1];
  in.bdf2 = c->hbdf[2];
  in.relax = cfg.heat_relaxation_factor;
  const double* al[3] = {c->alpha[0], c->alpha[1], c->alpha[2]};
  hg::LaunchHeatRows(c->st, c->geo, hb, al, c->np, cfg.conductivity, in, c->kcond, c->tgx,
                     c->tgy, c->As, c->Aw, c->Ac, c->Ae, c->An, c->k0);
  hg::Rows rows{c->As, c->Aw, c->Ac, c->Ae, c->An};
  switch (cfg.linear_solver_heat) {
    case HG_SOLVER_LU:
      hg::LaunchLuWavefront(c->st, nx, ny, rows, c->k0, nullptr, c->x0, nullptr, c->t_a,
                            nullptr, c->t_b, nullptr);
      break;
    case HG_SOLVER_LU_RB:
      // the red-black sweep takes two right-hand sides: the second is a copy
      hg::LaunchLuRb(c->st, nx, ny, rows, c->k0, c->k0, c->x0, c->x1, c->t_a, c->t_a, c->t_b,
                     c->tmp1);
      break;
    case HG_SOLVER_GAUSS_SEIDEL:
    case HG_SOLVER_JACOBI:
    case HG_SOLVER_LU_RELAXED: {
      int iters;
      double diff;
      SolveRowsIterative(c, cfg.linear_solver_heat, rows, c->k0, c->x0, &iters, &diff);
      KAddTo<<<G(c->nc), 256, 0, c->st>>>(c->nc, c->t_a, c->x0, c->t_b);
      break;
    }
    default:
      throw HgError(HG_ERR_UNSUPPORTED,
                    "linear_solver_heat: supported values are lu, lu_rb, gauss_seidel, jacobi, "
                    "lu_relaxed");
  }
  // time_prev <- time_curr <- iter_curr
  double* old_tp = c->t_tp;
  c->t_tp = c->t_tc;
  c->t_tc = c->t_b;
  c->t_b = c->t_a;
  c->t_a = old_tp;
  if (NanCheck(c, c->t_tc, c->t_tc, c->t_tc)) throw HgError(HG_ERR_NAN, "NaN field");
}

// The running part of CalcStat, run after every property update as the
// reference's step runs CalcStat after UpdateFluidProperties
// (hydro2d.hpp:1615-1617): the boundary in/out integrals over the
// time_curr volume flux (hydro2d.hpp:1474-1508) and the output offset
// meshpos (hydro2d.hpp:1526-1528).
void AccumulateStats(hg_ctx* c) {
  const double* al[3] = {c->alpha[0], c->alpha[1], c->alpha[2]};
  hg::LaunchBoundaryFlux(c->st, c->geo, c->fl_tc, al, c->np, c->cfg.density, c->dt,
                         c->bflux_part, c->bflux_acc);
  for (int d = 0; d < 2; ++d) c->meshpos_stat[d] = c->meshpos[d];
  if (c->cfg.meshvel_output)
    for (int d = 0; d < 2; ++d) c->meshpos[d] += c->cfg.meshvel[d] * c->dt;
}

// Constructor state of the running statistics, then the constructor's
// CalcStat (hydro2d.hpp:966-967).
void InitStats(hg_ctx* c) {
  HG_CHECK(hipMemsetAsync(c->bflux_acc, 0, 4 * HG_MAX_PHASES * sizeof(double), c->st));
  c->meshpos[0] = c->meshpos[1] = 0.;
  AccumulateStats(c);
}

void CalcStats(hg_ctx* c, hg_stats* s) {
  std::memset(s, 0, sizeof(*s));
  HG_CHECK(hipMemcpyAsync(c->stat_h + 8, c->bflux_acc, 4 * HG_MAX_PHASES * 8,
                          hipMemcpyDeviceToHost, c->st));
  for (int i = 0; i < c->np; ++i) {
    hg::LaunchStats(c->st, c->geo, c->xn_d, c->yn_d, c->alpha[i], c->pd[i], c->u_tc[0],
                    c->u_tc[1], c->stat_part, c->stat_out);
    HG_CHECK(hipMemcpyAsync(c->stat_h, c->stat_out, 7 * 8, hipMemcpyDeviceToHost, c->st));
    c->Sync();
    const double* v = c->stat_h;
    double volume = v[0];
    s->volume[i] = volume;
    s->mass[i] = volume * c->cfg.density[i];
    s->cx[i] = v[1] / volume + c->meshpos_stat[0];
    s->cy[i] = v[2] / volume + c->meshpos_stat[1];
    s->vx[i] = v[3] / volume;
    s->vy[i] = v[4] / volume;
    s->pd_min[i] = v[5];
    s->pd_max[i] = v[6];
    const double* acc = c->stat_h + 8 + 4 * i;
    s->volume_in[i] = acc[0];
    s->volume_out[i] = acc[1];
    s->mass_in[i] = acc[2];
    s->mass_out[i] = acc[3];
  }
}

// Time steps and the unsteady coefficients that follow from them
// (GetDerivativeApproxCoeffs at {-2dt, -dt, 0}, or {-dt, 0} for first order;
// conv_diff.hpp:201-203).
void SetTimeSteps(hg_ctx* c, double dt, double dt_adv) {
  c->dt = dt;
  c->dt_adv = dt_adv;
  std::vector<double> co = DerivCoeffs(0., {-2. * dt, -dt, 0.});
  if (!c->cfg.time_second_order) {
    std::vector<double> c2 = DerivCoeffs(0., {-dt, 0.});
    co = {0., c2[0], c2[1]};
  }
  for (int i = 0; i < 3; ++i) c->bdf[i] = co[i];
}

// dt_auto (hydro2d.hpp:1536-1544): dtm = FluidSimple::GetAutoTimeStep on the
// time_curr flux; fluid dt = dtm * cfl, advection dt = dtm * cfl_advection.
void AutoTimeStep(hg_ctx* c) {
  hg::LaunchAutoDt(c->st, c->geo, c->fl_tc, c->stat_part, c->stat_out);
  HG_CHECK(hipMemcpyAsync(c->stat_h, c->stat_out, 8, hipMemcpyDeviceToHost, c->st));
  c->Sync();
  const double dtm = c->stat_h[0];
  SetTimeSteps(c, dtm * c->cfg.cfl, dtm * c->cfg.cfl_advection);
}

// Fluid conditions per boundary face (hydro2d.hpp:370-392): walls and
// inlets hold their side's velocity, outlets start at zero
// (fluid.hpp:318-327); the inlet partial densities are the adjacent cells'
// (hydro2d.hpp:565-583).
void InitBoundary(hg_ctx* c) {
  const int nx = c->geo.nx, ny = c->geo.ny;
  const long nb = 2L * (nx + ny);
  std::vector<double> v(static_cast<size_t>(2 * nb), 0.);
  for (long b = 0; b < nb; ++b) {
    const int s = b < ny ? HG_SIDE_LEFT : b < 2 * ny ? HG_SIDE_RIGHT
                : b < 2 * ny + nx ? HG_SIDE_BOTTOM : HG_SIDE_TOP;
    if (c->side_kind[s] == HG_BC_OUTLET) continue;
    v[size_t(b)] = c->cfg.bc_velocity[s][0];
    v[size_t(nb + b)] = c->cfg.bc_velocity[s][1];
  }
  Upload(c, c->bvel_d, v);
  const double* pdv[3] = {c->pd[0], c->pd[1], c->pd[2]};
  hg::LaunchGatherInlet(c->st, c->geo, c->side_kind, pdv, c->np, c->pdb_d);
  c->Sync();  // v goes out of scope
}

void InitialFlux(hg_ctx* c) {
  hg::LaunchInitialFlux(c->st, c->geo, c->bc, c->u_tc[0], c->u_tc[1], c->cfg.meshvel[0],
                        c->cfg.meshvel[1], c->fl_tc);
  HG_CHECK(hipMemcpyAsync(c->fl_tp, c->fl_tc, c->nf * 8, hipMemcpyDeviceToDevice, c->st));
}

// Returns 0 when a, b and d are free of NaN / inf; else bit 1 for a, bit 2
// for b or d.
int NanCheck(hg_ctx* c, const double* a, const double* b, const double* d) {
  // the flag is zero between checks (zeroed at allocation, and again after
  // a check that set it), so no clearing pass precedes the kernel
  hg::LaunchNanCheck(c->st, a, b, d, int(c->nc), c->nan_d);
  HG_CHECK(hipMemcpyAsync(c->flag_h, c->nan_d, sizeof(int), hipMemcpyDeviceToHost, c->st));
  c->Sync();
  const int bits = *c->flag_h;
  if (bits == 0) return 0;
  HG_CHECK(hipMemsetAsync(c->nan_d, 0, sizeof(int), c->st));
  c->Sync();
  return bits;
}

void StartStep(hg_ctx* c) {
  c->fluid_iter = 0;
  c->cd_iter = 0;
  c->iter_rs.clear();
  // FluidSimple::StartStep checks the pressure, then the velocity solver's
  // StartStep the velocity (fluid.hpp:795-797, conv_diff.hpp:119-121)
  if (!c->fields_checked) {
    const int bits = NanCheck(c, c->p_tc, c->u_tc[0], c->u_tc[1]);
    if (bits & 1) throw HgError(HG_ERR_NAN, "NaN initial pressure");
    if (bits & 2) throw HgError(HG_ERR_NAN, "NaN initial field");
  }
  c->fields_checked = false;
  // iter_curr aliases time_curr after FinishStep: give it its own buffer
  if (c->p_ic == c->p_tc) c->p_ic = c->p_spare;
  if (c->fl_ic == c->fl_tc) c->fl_ic = c->fl_spare;
  for (int d = 0; d < 2; ++d)
    if (c->u_ic[d] == c->u_tc[d]) c->u_ic[d] = c->u_spare[d];
  if (c->guess_in_place) {
    // iter_curr = time_curr (the guess a + (a - b) 0 is a, up to the sign of
    // a zero); Iterate gives the iteration a buffer of its own
    c->p_ic = c->p_tc;
    c->fl_ic = c->fl_tc;
    for (int d = 0; d < 2; ++d) c->u_ic[d] = c->u_tc[d];
  } else {
    ExtrapLayers l{{c->u_tc[0], c->u_tc[1], c->p_tc, c->fl_tc},
                   {c->u_tp[0], c->u_tp[1], c->p_tp, c->fl_tp},
                   {c->u_ic[0], c->u_ic[1], c->p_ic, c->fl_ic}};
    KExtrapolate<<<G(c->nc + c->nf), 256, 0, c->st>>>(c->nc, c->nf, l,
                                                       c->cfg.guess_extrapolation);
  }
  c->flux32_valid = false;  // iter_curr's flux is the new guess
  // inputs that stay fixed through the SIMPLE loop (fluid.hpp:602-641)
  if (c->mixed) {
    auto& f = c->f32;
    hg::LaunchForceFaces(c->st, c->geo, c->fx, c->fy, c->mu_s, f.ffex, f.ffey, f.ffmu);
    hg::LaunchRestoredForce(c->st, c->geo, f.ffex, f.ffey, f.frx, f.fry);
    hg::LaunchFacesDeriv0Pair(c->st, c->geo, f.frx, f.fry, f.ffrx, f.ffry);
    return;
  }
  hg::LaunchForceFaces(c->st, c->geo, c->fx, c->fy, c->mu_s, c->ffex, c->ffey, c->ffmu);
  hg::LaunchRestoredForce(c->st, c->geo, c->ffex, c->ffey, c->frx, c->fry);
  hg::LaunchFacesDeriv0Pair(c->st, c->geo, c->frx, c->fry, c->ffrx, c->ffry);
}

// The reference's iterative row solvers from a zero guess into x:
// gauss_seidel (linear.hpp:685-715), jacobi (:750-783) and lu_relaxed
// (:592-648), all configured by the lu_relaxed_* keys (hydro2d.hpp:195-248).
// *iters and *diff receive the iteration count and final max|corr| the
// reference prints (linear.hpp:647,712,779).  Uses the Jacobi work arrays as
// scratch.
void SolveRowsIterative(hg_ctx* c, int32_t solver, hg::Rows rows, const double* k, double* x,
                        int* iters, double* diff) {
  const int nx = c->geo.nx, ny = c->geo.ny;
  const hg_config& cfg = c->cfg;
  const double tol = cfg.lu_relaxed_tolerance, omega = cfg.lu_relaxed_relaxation_factor;
  const int limit = cfg.lu_relaxed_num_iters_limit;
  hg::JacobiWork& w = c->jw;
  if (solver == HG_SOLVER_JACOBI) {
    HG_CHECK(hipMemsetAsync(w.xa, 0, c->nc * 8, c->st));
    HG_CHECK(hipMemsetAsync(w.diff, 0, size_t(limit + 2) * 8, c->st));
    HG_CHECK(hipMemsetAsync(w.state, 0, 4 * sizeof(int), c->st));
    int first = 0;
    while (first <= limit) {
      int n = std::min(64, limit + 1 - first);
      hg::LaunchJacobiSweeps(c->st, nx, ny, rows, k, w, first, n, tol, limit, omega);
      first += n;
      HG_CHECK(hipMemcpyAsync(c->flag_h, w.state, 4 * sizeof(int), hipMemcpyDeviceToHost, c->st));
      c->Sync();
      if (c->flag_h[0]) break;
    }
    const int res_buf = c->flag_h[3];
    const int it = c->flag_h[2];
    HG_CHECK(hipMemcpyAsync(x, res_buf ? w.xb : w.xa, c->nc * 8, hipMemcpyDeviceToDevice, c->st));
    unsigned long long bits;
    const int last_sweep = std::min(first, limit + 1) - 1;
    // the final diff is that of the last executed sweep
    HG_CHECK(hipMemcpyAsync(&bits, w.diff + std::max(0, std::min(it, last_sweep)), 8,
                            hipMemcpyDeviceToHost, c->st));
    c->Sync();
    std::memcpy(diff, &bits, 8);
    *iters = it;
    return;
  }
  if (solver == HG_SOLVER_GAUSS_SEIDEL)
    hg::LaunchGsWavefront(c->st, nx, ny, rows, k, x, tol, limit, omega, c->gs_info);
  else if (solver == HG_SOLVER_LU_RELAXED)
    hg::LaunchLuRelaxed(c->st, nx, ny, rows, k, x, w.xa, w.xb, tol, limit, omega, c->gs_info);
  else
    throw HgError(HG_ERR_UNSUPPORTED, "not an iterative row solver");
  HG_CHECK(hipMemcpyAsync(c->info_h, c->gs_info, 16, hipMemcpyDeviceToHost, c->st));
  c->Sync();
  *iters = int(c->info_h[0]);
  std::memcpy(diff, &c->info_h[1], 8);
}

inline bool IterativeRowSolver(int32_t s) {
  return s == HG_SOLVER_GAUSS_SEIDEL || s == HG_SOLVER_JACOBI || s == HG_SOLVER_LU_RELAXED;
}

void SolveVelocity(hg_ctx* c, hg::Rows rows, hg_iter_stats* is) {
  const int nx = c->geo.nx, ny = c->geo.ny;
  const hg_config& cfg = c->cfg;
  switch (cfg.linear_solver_velocity) {
    case HG_SOLVER_LU:
      hg::LaunchLuWavefront(c->st, nx, ny, rows, c->k0, c->k1, c->x0, c->x1, c->u_ip[0],
                            c->u_ip[1], c->u_ic[0], c->u_ic[1]);
      is->velocity_solve_iters[0] = is->velocity_solve_iters[1] = 0;
      break;
    case HG_SOLVER_LU_RB:
      hg::LaunchLuRb(c->st, nx, ny, rows, c->k0, c->k1, c->x0, c->x1, c->u_ip[0], c->u_ip[1],
                     c->u_ic[0], c->u_ic[1]);
      is->velocity_solve_iters[0] = is->velocity_solve_iters[1] = 0;
      break;
    case HG_SOLVER_GAUSS_SEIDEL:
    case HG_SOLVER_JACOBI:
    case HG_SOLVER_LU_RELAXED: {
      // one component after the other, as the reference solves them
      // (conv_diff.hpp:245, fluid.hpp:157-165)
      const double* kk[2] = {c->k0, c->k1};
      double* xx[2] = {c->x0, c->x1};
      for (int d = 0; d < 2; ++d) {
        double diff;
        SolveRowsIterative(c, cfg.linear_solver_velocity, rows, kk[d], xx[d],
                           &is->velocity_solve_iters[d], &diff);
        KAddTo<<<G(c->nc), 256, 0, c->st>>>(c->nc, c->u_ip[d], xx[d], c->u_ic[d]);
      }
      break;
    }
    default:
      throw HgError(HG_ERR_UNSUPPORTED,
                    "linear_solver_velocity: supported values are lu, lu_rb, gauss_seidel, "
                    "jacobi, lu_relaxed");
  }
}

void SolvePressure(hg_ctx* c, hg_iter_stats* is) {
  const int nx = c->geo.nx, ny = c->geo.ny;
  const hg_config& cfg = c->cfg;
  hg::Rows rows{c->Ps, c->Pw, c->Pc, c->Pe, c->Pn};
  is->pressure_converged = 1;
  switch (cfg.linear_solver_pressure) {
    case HG_SOLVER_GAUSS_SEIDEL:
    case HG_SOLVER_JACOBI:
    case HG_SOLVER_LU_RELAXED: {
      double d;
      SolveRowsIterative(c, cfg.linear_solver_pressure, rows, c->Pk, c->pcorr,
                         &is->pressure_solve_iters, &d);
      is->pressure_solve_diff = d;
      is->pressure_converged = !(d > cfg.lu_relaxed_tolerance);
      break;
    }
    case HG_SOLVER_LU:
      hg::LaunchLuWavefront(c->st, nx, ny, rows, c->Pk, nullptr, c->pcorr, nullptr, nullptr,
                            nullptr, nullptr, nullptr);
      is->pressure_solve_iters = 0;
      break;
    case HG_SOLVER_LU_RB:
      // one right-hand side: the second lane solves the same system into tmp1
      hg::LaunchLuRb(c->st, nx, ny, rows, c->Pk, c->Pk, c->x0, c->x1, c->zeros, c->zeros,
                     c->pcorr, c->tmp1);
      is->pressure_solve_iters = 0;
      break;
    case HG_SOLVER_CG:
    case HG_SOLVER_FCG_MG:
    case HG_SOLVER_FCG_LEARNED: {
      // operator k_f = -fcoef (negated on the way in); the right-hand side
      // pb = -(row constant) was written by the pressure right-hand side kernel
      c->cg->SetOperator(c->fcoef, c->fcoef + c->geo.nfx, -1.);
      hg::Precond* pc = nullptr;
      if (cfg.linear_solver_pressure != HG_SOLVER_CG) {
        c->mg->Build(c->st);
        const bool learned = cfg.linear_solver_pressure == HG_SOLVER_FCG_LEARNED && c->net &&
                             c->net->Trained();
        pc = learned ? static_cast<hg::Precond*>(c->net.get())
                     : static_cast<hg::Precond*>(c->mg.get());
        if (cfg.fcg_negate_precond_gpu) {
          c->negated.reset(new NegatedPrecond(pc, c->nc));
          pc = c->negated.get();
        }
      }
      c->pressure_system = true;
      // the closing residual check is read back at the next synchronisation;
      // the first SIMPLE iteration of a step may start from the previous
      // step's correction (see hg_ctx::pcorr_warm)
      const bool warm = cfg.pressure_warm_start_gpu && c->pcorr_warm && c->fluid_iter == 0;
      c->pcorr_converged = false;
      c->cg_res = c->cg->Solve(c->pb, c->pcorr, cfg.pressure_tolerance_gpu,
                               cfg.pressure_max_iters_gpu, cfg.precision_gpu, true, pc, true,
                               warm);
      c->pending_cg = true;
      is->pressure_solve_iters = c->cg_res.iterations;
      break;
    }
    default:
      throw HgError(HG_ERR_UNSUPPORTED, "linear_solver_pressure value not supported");
  }
}

// The FP64 path's part of a SIMPLE iteration up to the fluxes: cell
// gradients, momentum rows (fluid.hpp:827-874, conv_diff.hpp:130-243), the
// velocity solve and Rhie-Chow (fluid.hpp:876-970).
void Fp64Momentum(hg_ctx* c, hg_iter_stats* is) {
  const hg_config& cfg = c->cfg;
  hg::LaunchCellGrads(c->st, c->geo, c->bc, c->p_ip, c->u_ip[0], c->u_ip[1], c->gpx, c->gpy,
                      c->g0x, c->g0y, c->g1x, c->g1y);
  hg::MomentumArgs m{};
  m.u0 = c->u_ip[0];
  m.u1 = c->u_ip[1];
  m.g0x = c->g0x; m.g0y = c->g0y; m.g1x = c->g1x; m.g1y = c->g1y;
  m.gpx = c->gpx; m.gpy = c->gpy; m.frx = c->frx; m.fry = c->fry;
  m.ffmu = c->ffmu;
  m.flux = c->fl_ip;
  m.rho = c->rho_s;
  m.utp0 = c->u_tp[0]; m.utp1 = c->u_tp[1]; m.utc0 = c->u_tc[0]; m.utc1 = c->u_tc[1];
  m.bdf0 = c->bdf[0]; m.bdf1 = c->bdf[1]; m.bdf2 = c->bdf[2];
  m.relax = cfg.velocity_relaxation_factor;
  m.s = c->As; m.w = c->Aw; m.c = c->Ac; m.e = c->Ae; m.n = c->An; m.k0 = c->k0; m.k1 = c->k1;
  m.ap = c->ap;  // the diagonal of fluid.hpp:876-883, from the same rows
  hg::LaunchMomentum(c->st, c->geo, c->bc, m);
  hg::Rows vrows{c->As, c->Aw, c->Ac, c->Ae, c->An};
  SolveVelocity(c, vrows, is);
  c->cd_iter += 1;
  hg::LaunchRhieChow(c->st, c->geo, c->bc, c->u_ic[0], c->u_ic[1], c->gpx, c->gpy, c->ffrx,
                     c->ffry, c->ffex, c->ffey, c->p_ip, c->ap, cfg.rhie_chow_factor,
                     cfg.meshvel[0], cfg.meshvel[1], c->fstar, c->fcoef);
}

// The mixed path's part of a SIMPLE iteration up to the fluxes: FP32 cell
// gradients (differences taken in FP64), momentum rows with FP64 residual
// constants stored as FP32, the red-black velocity correction in FP32 added
// to the FP64 iterate, and Rhie-Chow in FP64 on the FP32 intermediates.
void MixedMomentum(hg_ctx* c, hg_iter_stats* is) {
  const hg_config& cfg = c->cfg;
  auto& f = c->f32;
  hg::LaunchCellGrads(c->st, c->geo, c->bc, c->p_ip, c->u_ip[0], c->u_ip[1], f.gpx, f.gpy,
                      f.g0x, f.g0y, f.g1x, f.g1y);
  hg::MomentumArgsF m{};
  m.u0 = c->u_ip[0];
  m.u1 = c->u_ip[1];
  m.g0x = f.g0x; m.g0y = f.g0y; m.g1x = f.g1x; m.g1y = f.g1y;
  m.gpx = f.gpx; m.gpy = f.gpy; m.frx = f.frx; m.fry = f.fry;
  m.ffmu = f.ffmu;
  m.flux = c->fl_ip;
  m.rho = c->rho_s;
  m.utp0 = c->u_tp[0]; m.utp1 = c->u_tp[1]; m.utc0 = c->u_tc[0]; m.utc1 = c->u_tc[1];
  m.bdf0 = c->bdf[0]; m.bdf1 = c->bdf[1]; m.bdf2 = c->bdf[2];
  m.relax = cfg.velocity_relaxation_factor;
  m.s = f.s; m.w = f.w; m.c = f.c; m.e = f.e; m.n = f.n; m.k0 = f.k0; m.k1 = f.k1;
  m.ap = c->ap;
  hg::LaunchMomentum(c->st, c->geo, c->bc, m);
  hg::RowsF rows{f.s, f.w, f.c, f.e, f.n};
  hg::LaunchLuRb(c->st, c->geo.nx, c->geo.ny, rows, f.k0, f.k1, f.xb0, f.xb1, c->u_ip[0],
                 c->u_ip[1], c->u_ic[0], c->u_ic[1]);
  is->velocity_solve_iters[0] = is->velocity_solve_iters[1] = 0;
  c->cd_iter += 1;
  hg::LaunchRhieChow(c->st, c->geo, c->bc, c->u_ic[0], c->u_ic[1], f.gpx, f.gpy, f.ffrx, f.ffry,
                     f.ffex, f.ffey, c->p_ip, c->ap, cfg.rhie_chow_factor, cfg.meshvel[0],
                     cfg.meshvel[1], c->fstar, c->fcoef);
}

// One SIMPLE iteration; its statistics are built in c->last_iter and copied
// to *out.  With defer_rs the convergence indicator (and the pressure
// solve's closing check) are left for the next Sync(): for callers whose
// loop ends on the iteration limit without looking at the indicator.
void Iterate(hg_ctx* c, hg_iter_stats* out, bool defer_rs = false) {
  hg_iter_stats* is = &c->last_iter;
  std::memset(is, 0, sizeof(*is));
  const hg_config& cfg = c->cfg;
  // outlet velocities from the current iterate (fluid.hpp:820-821)
  if (c->any_outlet) {
    const long nb = 2L * (c->geo.nx + c->geo.ny);
    hg::LaunchOutletUpdate(c->st, c->geo, c->side_kind, c->u_ic[0], c->u_ic[1], c->bvel_d,
                           c->bvel_d + nb, c->obc_scratch);
  }
  // iter_prev <- iter_curr (fluid.hpp:815-818, conv_diff.hpp:132-134); the
  // iter_curr buffers are rewritten in full below.
  std::swap(c->p_ip, c->p_ic);
  std::swap(c->fl_ip, c->fl_ic);
  std::swap(c->u_ip[0], c->u_ic[0]);
  std::swap(c->u_ip[1], c->u_ic[1]);
  if (c->guess_in_place) {
    // with the guess read in place, iter_curr can now name a time layer:
    // write the iteration into the buffer no other layer holds
    auto own = [](double*& ic, double* tp, double* tc, double* ip, double* const bufs[4]) {
      if (ic != tp && ic != tc && ic != ip) return;
      for (int k = 0; k < 4; ++k)
        if (bufs[k] != tp && bufs[k] != tc && bufs[k] != ip) {
          ic = bufs[k];
          return;
        }
    };
    own(c->u_ic[0], c->u_tp[0], c->u_tc[0], c->u_ip[0], c->layer_bufs[0]);
    own(c->u_ic[1], c->u_tp[1], c->u_tc[1], c->u_ip[1], c->layer_bufs[1]);
    own(c->p_ic, c->p_tp, c->p_tc, c->p_ip, c->layer_bufs[2]);
    own(c->fl_ic, c->fl_tp, c->fl_tc, c->fl_ip, c->layer_bufs[3]);
  }
  if (c->mixed)
    MixedMomentum(c, is);
  else
    Fp64Momentum(c, is);
  if (FaceSolver(cfg.linear_solver_pressure))
    hg::LaunchPressureRhs(c->st, c->geo, c->fstar, c->pb, true);  // operator from the faces
  else
    hg::LaunchPressureRows(c->st, c->geo, c->fstar, c->fcoef, c->pinned,
                           cfg.pressure_fixed_value, c->Ps, c->Pw, c->Pc, c->Pe, c->Pn, c->Pk);
  SolvePressure(c, is);
  hg::LaunchCorrectCells(c->st, c->geo, c->pcorr, c->ap, c->p_ip,
                         cfg.pressure_relaxation_factor, c->p_ic, c->u_ic[0], c->u_ic[1],
                         c->u_ip[0], c->u_ip[1], c->rs_part, (double*)c->rs_d);
  hg::LaunchCorrectFaces(c->st, c->geo, c->fstar, c->fcoef, c->pcorr, c->fl_ic,
                         c->mixed ? c->f32.flux : nullptr);
  c->flux32_valid = c->mixed;
  if (cfg.simpler) {
    // SIMPLER (fluid.hpp:1059-1151): a second solve on the same pressure
    // rows with the right-hand side of KSimplerRhs, added to p; the
    // iteration's velocity gradients are dead by now and hold the scratch
    hg::Rows vrows2{c->As, c->Aw, c->Ac, c->Ae, c->An};
    hg::Rows prow{c->Ps, c->Pw, c->Pc, c->Pe, c->Pn};
    hg::LaunchSimpler(c->st, c->geo, vrows2, c->k0, c->k1, c->u_ic[0], c->u_ic[1], c->u_ip[0],
                      c->u_ip[1], c->frx, c->fry, c->gpx, c->gpy, c->ffex, c->ffey, c->ap,
                      c->fl_ic, cfg.rhie_chow_factor, prow, c->p_ic, c->pinned,
                      cfg.pressure_fixed_value, c->g0x, c->g0y, c->Pk);
    const hg_iter_stats first = *is;
    SolvePressure(c, is);
    KAddInPlace<<<G(c->nc), 256, 0, c->st>>>(c->nc, c->pcorr, c->p_ic);
    is->pressure_solve_iters += first.pressure_solve_iters;
    is->pressure_solve_diff = first.pressure_solve_diff;
    is->pressure_converged = first.pressure_converged && is->pressure_converged;
  }
  HG_CHECK(hipMemcpyAsync(c->rs_h, c->rs_d, 8, hipMemcpyDeviceToHost, c->st));
  c->pending_rs = true;
  c->fluid_iter += 1;
  c->last_s = c->fluid_iter;
  is->iteration = c->fluid_iter;
  c->iter_rs.push_back(0.);  // set when the indicator settles
  if (!defer_rs) c->Sync();  // settles the indicator into last_rs / last_iter
  *out = *is;
}

bool IsConverged(const hg_ctx* c) {
  double ind = (c->cd_iter == 0) ? 1. : c->last_rs;
  return c->fluid_iter >= c->cfg.num_iterations_limit || ind < c->cfg.convergence_tolerance;
}

// Layer shift time_prev <- time_curr <- iter_curr (fluid.hpp:1159-1163) by
// pointer rotation: iter_curr's buffer becomes time_curr, and iter_curr
// keeps pointing at it (the two layers hold the same values until the next
// StartStep, which gives iter_curr the freed time_prev buffer back).
void Rotate(double*& tp, double*& tc, double*& ic, double*& spare) {
  spare = tp;
  tp = tc;
  tc = ic;
}

void FinishStep(hg_ctx* c) {
  Rotate(c->p_tp, c->p_tc, c->p_ic, c->p_spare);
  Rotate(c->fl_tp, c->fl_tc, c->fl_ic, c->fl_spare);
  for (int d = 0; d < 2; ++d) Rotate(c->u_tp[d], c->u_tc[d], c->u_ic[d], c->u_spare[d]);
  // the checks come before the time advances, as FinishStep throws before
  // IncTime (fluid.hpp:1164-1168, conv_diff.hpp:255-258): pressure first
  const int bits = NanCheck(c, c->p_tc, c->u_tc[0], c->u_tc[1]);
  if (bits & 1) throw HgError(HG_ERR_NAN, "NaN pressure");
  if (bits & 2) throw HgError(HG_ERR_NAN, "NaN field");
  c->t_fluid += c->dt;
  c->steps += 1;
  c->mod_t = c->mod_tnp;
  c->mod_tnp = c->mod_t + c->dt;
  c->fields_checked = true;
  // (NanCheck synchronised: the step's last pressure solve is settled)
  c->pcorr_warm = c->fluid_iter == 1 && c->pcorr_converged;
}

void AdvectSubstep(hg_ctx* c) {
  // all phase fields in one pass (the reference advects them one after the
  // other with the same flux; the fields do not interact within a sub-step)
  const double* u[3] = {c->pd[0], c->np > 1 ? c->pd[1] : nullptr, c->np > 2 ? c->pd[2] : nullptr};
  double* un[3] = {c->pd_new[0], c->np > 1 ? c->pd_new[1] : nullptr,
                   c->np > 2 ? c->pd_new[2] : nullptr};
  const long nb = 2L * (c->geo.nx + c->geo.ny);
  const double* bval[3] = {nullptr, nullptr, nullptr};
  if (c->any_inlet)
    for (int i = 0; i < c->np; ++i) bval[i] = c->pdb_d + i * nb;
  int inlet[4];
  for (int s = 0; s < 4; ++s) inlet[s] = c->side_kind[s] == HG_BC_INLET ? 1 : 0;
  if (c->mixed) {
    if (!c->flux32_valid) {
      hg::LaunchToFloat(c->st, c->nf, c->fl_ic, c->f32.flux);
      c->flux32_valid = true;
    }
    hg::LaunchAdvStep(c->st, c->geo, c->np, u, c->f32.flux, c->dt_adv, un,
                      c->any_inlet ? bval : nullptr, inlet);
  } else {
    hg::LaunchAdvStep(c->st, c->geo, c->np, u, c->fl_ic, c->dt_adv, un,
                      c->any_inlet ? bval : nullptr, inlet);
  }
  for (int i = 0; i < c->np; ++i) std::swap(c->pd[i], c->pd_new[i]);
  c->t_adv += c->dt_adv;
}

hg_status Guard(hg_ctx* ctx, const std::function<void()>& fn) {
  try {
    fn();
    return HG_OK;
  } catch (const HgError& e) {
    if (ctx) {
      ctx->err = e.what();
      ctx->pending_cg = ctx->pending_rs = false;  // a failed call settles nothing
    }
    return e.code;
  } catch (const std::exception& e) {
    if (ctx) {
      ctx->err = e.what();
      ctx->pending_cg = ctx->pending_rs = false;
    }
    return HG_ERR_HIP;
  }
}

}  // namespace

namespace hg {
// hg_last_error(NULL) reports the message of the last failed context-free
// call on this thread.
void SetThreadError(const std::string& msg);
}

// ===========================================================================
extern "C" {

// Message of the last failed hg_create (or context-free call) on this thread.
static thread_local std::string g_create_error;

hg_status hg_create(const hg_config* cfg_in, hg_ctx** out) {
  if (!cfg_in || !out) return HG_ERR_BAD_ARG;
  *out = nullptr;
  const hg_config& cfg = *cfg_in;
  if (cfg.abi_version != HG_ABI_VERSION) return HG_ERR_BAD_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return HG_ERR_HIP;
  std::unique_ptr<hg_ctx> c(new hg_ctx());
  c->cfg = cfg;
  hg_status rc = Guard(c.get(), [&]() {
    if (cfg.nx < 2 || cfg.ny < 2) throw HgError(HG_ERR_BAD_ARG, "Nx, Ny must be >= 2");
    if (cfg.num_phases < 1 || cfg.num_phases > 3) throw HgError(HG_ERR_BAD_ARG, "num_phases 1..3");
    if (cfg.unsupported_flags)
      throw HgError(HG_ERR_UNSUPPORTED, "script enables features outside this path (flags " +
                                            std::to_string(cfg.unsupported_flags) + ")");
    if (cfg.force_geometric_average)
      throw HgError(HG_ERR_UNSUPPORTED, "force_geometric_average is not supported");
    // solver names are checked here, before any step runs (the reference's
    // factory throws at construction, hydro2d.hpp:247, 267-276)
    {
      const int32_t v = cfg.linear_solver_velocity, p = cfg.linear_solver_pressure,
                    h = cfg.linear_solver_heat;
      const bool row_solver_v = v == HG_SOLVER_LU || v == HG_SOLVER_LU_RB ||
                                v == HG_SOLVER_GAUSS_SEIDEL || v == HG_SOLVER_JACOBI ||
                                v == HG_SOLVER_LU_RELAXED;
      if (!row_solver_v)
        throw HgError(HG_ERR_UNSUPPORTED,
                      "linear_solver_velocity: supported values are lu, lu_rb, gauss_seidel, "
                      "jacobi, lu_relaxed");
      if (p < HG_SOLVER_LU || p > HG_SOLVER_FCG_MG)
        throw HgError(HG_ERR_BAD_ARG, "linear_solver_pressure: unknown solver");
      if (cfg.heat_enable && !(h == HG_SOLVER_LU || h == HG_SOLVER_LU_RB ||
                               h == HG_SOLVER_GAUSS_SEIDEL || h == HG_SOLVER_JACOBI ||
                               h == HG_SOLVER_LU_RELAXED))
        throw HgError(HG_ERR_UNSUPPORTED,
                      "linear_solver_heat: supported values are lu, lu_rb, gauss_seidel, "
                      "jacobi, lu_relaxed");
      if (FaceSolver(p) && cfg.has_pressure_fixed_point)
        throw HgError(HG_ERR_UNSUPPORTED,
                      "cg / fcg solve the pressure up to a constant: pressure_fixed_point is not "
                      "supported with them");
    }
    if (cfg.precision_gpu == HG_PREC_MIXED) {
      if (cfg.linear_solver_velocity != HG_SOLVER_LU_RB)
        throw HgError(HG_ERR_UNSUPPORTED,
                      "precision_gpu mixed solves the velocity correction in FP32 with lu_rb: "
                      "set linear_solver_velocity lu_rb");
      if (cfg.simpler)
        throw HgError(HG_ERR_UNSUPPORTED, "simpler runs in the FP64 path only");
    }
    if (cfg.simpler && FaceSolver(cfg.linear_solver_pressure))
      throw HgError(HG_ERR_UNSUPPORTED,
                    "simpler runs with the row solvers (gauss_seidel, jacobi, lu), not cg / fcg");
    if (cfg.sigma != 0. && cfg.num_phases >= 2)
      throw HgError(HG_ERR_UNSUPPORTED, "sigma != 0 is not supported");
    if (std::abs(cfg.sharp) > 1e-10) throw HgError(HG_ERR_UNSUPPORTED, "sharp != 0 is not supported");
    if (cfg.tvd_split) throw HgError(HG_ERR_UNSUPPORTED, "tvd_split is not supported");
    if (long(cfg.nx) * cfg.ny > (1L << 30)) throw HgError(HG_ERR_BAD_ARG, "mesh too large");
    HG_CHECK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    for (auto& e : c->ev) HG_CHECK(hipEventCreate(&e));
    for (auto& e : c->kev) HG_CHECK(hipEventCreate(&e));
    c->np = cfg.num_phases;
    c->mesh.Init(cfg.nx, cfg.ny, cfg.a, cfg.b);
    c->geo = c->mesh.geo;
    for (int d = 0; d < 2; ++d)
      for (int s = 0; s < 4; ++s) c->bc.vel[d][s] = cfg.bc_velocity[s][d];
    for (int s = 0; s < 4; ++s) {
      c->side_kind[s] = cfg.bc_kind[s];
      c->any_inlet = c->any_inlet || cfg.bc_kind[s] == HG_BC_INLET;
      c->any_outlet = c->any_outlet || cfg.bc_kind[s] == HG_BC_OUTLET;
    }
    c->open_bc = c->any_inlet || c->any_outlet;
    c->ph.n = c->np;
    for (int i = 0; i < 3; ++i) {
      c->ph.rho[i] = cfg.density[i];
      c->ph.mu[i] = cfg.viscosity[i];
    }
    for (int d = 0; d < 2; ++d) {
      c->ph.grav[d] = cfg.gravity[d];
      c->ph.force[d] = cfg.force[d];
    }
    SetTimeSteps(c.get(), cfg.dt, cfg.dt * cfg.advection_dt_factor);
    c->mod_tnp = 0. + cfg.dt;
    const long nc = long(cfg.nx) * cfg.ny;
    const long nf = long(cfg.nx + 1) * cfg.ny + long(cfg.nx) * (cfg.ny + 1);
    c->nc = nc;
    c->nf = nf;
    for (int d = 0; d < 2; ++d) {
      c->u_tc[d] = c->Alloc(nc); c->u_tp[d] = c->Alloc(nc);
      c->u_ic[d] = c->Alloc(nc); c->u_ip[d] = c->Alloc(nc);
    }
    c->p_tc = c->Alloc(nc); c->p_tp = c->Alloc(nc); c->p_ic = c->Alloc(nc); c->p_ip = c->Alloc(nc);
    c->fl_tc = c->Alloc(nf); c->fl_tp = c->Alloc(nf); c->fl_ic = c->Alloc(nf); c->fl_ip = c->Alloc(nf);
    for (int d = 0; d < 2; ++d) {
      double* b[4] = {c->u_tc[d], c->u_tp[d], c->u_ic[d], c->u_ip[d]};
      std::copy(b, b + 4, c->layer_bufs[d]);
    }
    {
      double* bp[4] = {c->p_tc, c->p_tp, c->p_ic, c->p_ip};
      double* bf[4] = {c->fl_tc, c->fl_tp, c->fl_ic, c->fl_ip};
      std::copy(bp, bp + 4, c->layer_bufs[2]);
      std::copy(bf, bf + 4, c->layer_bufs[3]);
    }
    c->guess_in_place = cfg.precision_gpu == HG_PREC_MIXED && cfg.guess_extrapolation == 0. &&
                        cfg.fluid_enable;
    for (double** p : {&c->rho, &c->rho_s, &c->mu, &c->mu_s, &c->fx, &c->fy, &c->tmp1, &c->tmp2})
      *p = c->Alloc(nc);
    for (int i = 0; i < 3; ++i) {
      c->pd[i] = c->Alloc(i < c->np ? nc : 1);
      c->pd_new[i] = c->Alloc(i < c->np ? nc : 1);
      c->alpha[i] = c->Alloc(i < c->np ? nc : 1);
    }
    for (double** p : {&c->ffex, &c->ffey, &c->ffmu, &c->ffrx, &c->ffry, &c->fstar, &c->fcoef,
                       &c->kface})
      *p = c->Alloc(nf);
    for (double** p : {&c->frx, &c->fry, &c->gpx, &c->gpy, &c->g0x, &c->g0y, &c->g1x, &c->g1y,
                       &c->As, &c->Aw, &c->Ac, &c->Ae, &c->An, &c->k0, &c->k1, &c->x0, &c->x1,
                       &c->ap, &c->Ps, &c->Pw, &c->Pc, &c->Pe, &c->Pn, &c->Pk, &c->pcorr, &c->pb,
                       &c->zeros})
      *p = c->Alloc(nc);
    if (cfg.precision_gpu == HG_PREC_MIXED) {
      c->mixed = true;
      auto& f = c->f32;
      for (float** p : {&f.ffex, &f.ffey, &f.ffmu, &f.ffrx, &f.ffry, &f.flux}) *p = c->AllocF(nf);
      for (float** p : {&f.frx, &f.fry, &f.gpx, &f.gpy, &f.g0x, &f.g0y, &f.g1x, &f.g1y, &f.s,
                        &f.w, &f.c, &f.e, &f.n, &f.k0, &f.k1, &f.xb0, &f.xb1})
        *p = c->AllocF(nc);
    }
    c->xn_d = c->Alloc(cfg.nx + 1);
    c->yn_d = c->Alloc(cfg.ny + 1);
    c->stat_part = c->Alloc(long(hg::StatsBlocks(c->geo)) * 7);
    c->stat_out = c->Alloc(8);
    c->bflux_acc = c->Alloc(4 * HG_MAX_PHASES);
    c->bflux_part = c->Alloc(6 * long(hg::BoundaryFluxBlocks(c->geo)));
    if (cfg.heat_enable) InitHeat(c.get());
    c->rs_part = c->Alloc((nc + 255) / 256);
    c->rs_d = (unsigned long long*)c->Alloc(1);
    c->nan_d = (int*)c->Alloc(1);
    c->gs_info = (unsigned long long*)c->Alloc(2);
    c->jw.xa = c->Alloc(nc);
    c->jw.xb = c->Alloc(nc);
    c->jw.diff = (unsigned long long*)c->Alloc(cfg.lu_relaxed_num_iters_limit + 2);
    c->jw.state = (int*)c->Alloc(2);
    HG_CHECK(hipHostMalloc((void**)&c->rs_h, 8));
    HG_CHECK(hipHostMalloc((void**)&c->flag_h, 4 * sizeof(int)));
    HG_CHECK(hipHostMalloc((void**)&c->info_h, 16));
    HG_CHECK(hipHostMalloc((void**)&c->stat_h, (8 + 4 * HG_MAX_PHASES) * 8));
    const int32_t lsp = cfg.linear_solver_pressure;
    if (lsp == HG_SOLVER_CG || lsp == HG_SOLVER_FCG_LEARNED || lsp == HG_SOLVER_FCG_MG)
      c->cg.reset(new hg::CgSolver(cfg.nx, cfg.ny, c->st));
    if (lsp == HG_SOLVER_FCG_LEARNED || lsp == HG_SOLVER_FCG_MG)
      c->mg.reset(new hg::MgPrecond(cfg.nx, cfg.ny, c->cg->kx32(), c->cg->ky32(),
                                    c->cg->dinv32()));
    Upload(c.get(), c->xn_d, c->mesh.xn);
    Upload(c.get(), c->yn_d, c->mesh.yn);

    // ---- initial state (hydro ctor: InitFluidSolver, InitAdvectionSolver,
    //      UpdateFluidProperties; hydro2d.hpp:267-615, 928-972)
    std::vector<double> cx(nc), cy(nc);
    for (int j = 0; j < cfg.ny; ++j)
      for (int i = 0; i < cfg.nx; ++i) c->mesh.CellCenter(i, j, cx[i + long(cfg.nx) * j],
                                                          cy[i + long(cfg.nx) * j]);
    std::vector<double> u0(nc, cfg.initial_velocity[0]), u1(nc, cfg.initial_velocity[1]);
    if (cfg.has_initial_sin) {
      double pi = std::atan(1.) * 4.;
      double kx = cfg.initial_sin_n[0] * (2. * pi / cfg.initial_sin_lambda);
      double ky = cfg.initial_sin_n[1] * (2. * pi / cfg.initial_sin_lambda);
      for (long q = 0; q < nc; ++q) {
        double dotv = 0.;
        dotv += kx * cx[q];
        dotv += ky * cy[q];
        double sv = std::sin(dotv - cfg.initial_sin_phase);
        u0[q] *= sv;
        u1[q] *= sv;
      }
    }
    Upload(c.get(), c->u_tc[0], u0);
    Upload(c.get(), c->u_tc[1], u1);
    Upload(c.get(), c->u_tp[0], u0);
    Upload(c.get(), c->u_tp[1], u1);
    if (cfg.has_pressure_fixed_point) {
      auto dist = [&](long q) {
        double dx = cx[q] - cfg.pressure_fixed_point[0], dy = cy[q] - cfg.pressure_fixed_point[1];
        double s = 0.;
        s += dx * dx;
        s += dy * dy;
        return std::sqrt(s);
      };
      long best = 0;
      for (long q = 0; q < nc; ++q)
        if (dist(q) < dist(best)) best = q;
      c->pinned = int(best);
    }
    // partial densities before smoothing
    std::vector<std::vector<double>> pdh(c->np);
    for (int i = 0; i < c->np; ++i)
      pdh[i].assign(nc, cfg.density[i] * cfg.initial_volume_fraction[i]);
    for (long q = 0; q < nc; ++q) {
      double x = cx[q], y = cy[q];
      auto inside = [&](const double* lo, const double* hi) {
        return !(x < lo[0] || hi[0] < x) && !(y < lo[1] || hi[1] < y);
      };
      auto dist2 = [&](const double* ctr) {
        double dx = ctr[0] - x, dy = ctr[1] - y, s = 0.;
        s += dx * dx;
        s += dy * dy;
        return std::sqrt(s);
      };
      if (inside(cfg.a2, cfg.b2)) {
        if (c->np < 3) throw HgError(HG_ERR_BAD_ARG, "block A2-B2 holds cells but num_phases < 3");
        pdh[2][q] = cfg.density[2];
      } else if (c->np >= 2 && inside(cfg.a1, cfg.b1)) {
        pdh[1][q] = cfg.density[1];
      } else if (c->np >= 2 && dist2(cfg.ic) < cfg.ir) {
        pdh[1][q] = cfg.density[1];
      } else if (c->np >= 2 && dist2(cfg.ic2) < cfg.ir2) {
        pdh[1][q] = cfg.density[1];
      }
    }
    for (int i = 0; i < c->np; ++i) Upload(c.get(), c->pd[i], pdh[i]);
    for (int i = 1; i < c->np; ++i)
      Smooth(c.get(), c->pd[i], c->pd[i], cfg.initial_volume_fraction_smooth_times);
    hg::LaunchPhase0Init(c->st, c->geo, c->ph, c->pd[1], c->pd[2], c->pd[0]);
    if (c->open_bc) {
      const long nb = 2L * (cfg.nx + cfg.ny);
      c->bvel_d = c->Alloc(2 * nb);
      c->pdb_d = c->Alloc(c->np * nb);
      c->obc_scratch = c->Alloc(2 * nb + 1);
      c->bc.fvel[0] = c->bvel_d;
      c->bc.fvel[1] = c->bvel_d + nb;
      InitBoundary(c.get());
    }
    InitialFlux(c.get());
    UpdateProperties(c.get());
    InitStats(c.get());
    c->Sync();
  });
  if (rc != HG_OK) {
    // keep the message reachable: hand back the context only on success
    g_create_error = c->err;
    std::fprintf(stderr, "hg_create: %s\n", g_create_error.c_str());
    return rc;
  }
  *out = c.release();
  return HG_OK;
}

void hg_destroy(hg_ctx* ctx) { delete ctx; }

}  // extern "C"

void hg::SetThreadError(const std::string& msg) { g_create_error = msg; }

extern "C" {

const char* hg_last_error(const hg_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_create_error.c_str();
}
int32_t hg_num_cells(const hg_ctx* ctx) { return ctx ? int32_t(ctx->nc) : 0; }
int32_t hg_num_faces(const hg_ctx* ctx) { return ctx ? int32_t(ctx->nf) : 0; }

static double* FieldPtr(hg_ctx* c, int f, bool write, long* n) {
  *n = c->nc;
  switch (f) {
    case HG_FIELD_VELOCITY_X: return c->u_tc[0];
    case HG_FIELD_VELOCITY_Y: return c->u_tc[1];
    case HG_FIELD_PRESSURE: return c->p_tc;
    case HG_FIELD_VOLUME_FLUX: *n = c->nf; return c->fl_tc;
    case HG_FIELD_TEMPERATURE: return c->t_tc;  // nullptr unless heat_enable
    default: break;
  }
  if (f >= HG_FIELD_PARTIAL_DENSITY_0 && f < HG_FIELD_PARTIAL_DENSITY_0 + c->np)
    return c->pd[f - HG_FIELD_PARTIAL_DENSITY_0];
  if (write) return nullptr;
  if (f >= HG_FIELD_VOLUME_FRACTION_0 && f < HG_FIELD_VOLUME_FRACTION_0 + c->np)
    return c->alpha[f - HG_FIELD_VOLUME_FRACTION_0];
  switch (f) {
    case HG_FIELD_DENSITY: return c->rho;
    case HG_FIELD_DENSITY_SMOOTH: return c->rho_s;
    case HG_FIELD_VISCOSITY_SMOOTH: return c->mu_s;
    case HG_FIELD_FORCE_X: return c->fx;
    case HG_FIELD_FORCE_Y: return c->fy;
    case HG_FIELD_VOLUME_FLUX_ITER: *n = c->nf; return c->fl_ic;
    case HG_FIELD_PRESSURE_FACE_COEF:
      *n = c->nf;
      KNegateFaces<<<G(c->nf), 256, 0, c->st>>>(c->nf, c->fcoef, c->kface);
      return c->kface;
    case HG_FIELD_PRESSURE_RHS:
      // the face-form solvers' right-hand side is written negated already
      if (!FaceSolver(c->cfg.linear_solver_pressure))
        KNeg<<<G(c->nc), 256, 0, c->st>>>(c->nc, c->Pk, c->pb);
      return c->pb;
    default: return nullptr;
  }
}

hg_status hg_set_field(hg_ctx* c, int32_t field, const double* host, size_t n) {
  if (!c || !host) return HG_ERR_BAD_ARG;
  long len;
  double* d = FieldPtr(c, field, true, &len);
  if (!d || size_t(len) != n) {
    c->err = "hg_set_field: bad field or size";
    return HG_ERR_BAD_ARG;
  }
  return Guard(c, [&]() {
    c->fields_checked = false;
    c->flux32_valid = false;
    HG_CHECK(hipMemcpyAsync(d, host, n * 8, hipMemcpyHostToDevice, c->st));
    double* prev = nullptr;
    if (field == HG_FIELD_VELOCITY_X) prev = c->u_tp[0];
    if (field == HG_FIELD_VELOCITY_Y) prev = c->u_tp[1];
    if (field == HG_FIELD_PRESSURE) prev = c->p_tp;
    if (field == HG_FIELD_VOLUME_FLUX) prev = c->fl_tp;
    if (field == HG_FIELD_TEMPERATURE) prev = c->t_tp;
    if (prev) HG_CHECK(hipMemcpyAsync(prev, host, n * 8, hipMemcpyHostToDevice, c->st));
    c->Sync();
  });
}

hg_status hg_get_field(hg_ctx* c, int32_t field, double* host, size_t n) {
  if (!c || !host) return HG_ERR_BAD_ARG;
  long len;
  double* d = FieldPtr(c, field, false, &len);
  if (!d || size_t(len) != n) {
    c->err = "hg_get_field: bad field or size";
    return HG_ERR_BAD_ARG;
  }
  return Guard(c, [&]() {
    HG_CHECK(hipMemcpyAsync(host, d, n * 8, hipMemcpyDeviceToHost, c->st));
    c->Sync();
  });
}

hg_status hg_get_config(const hg_ctx* c, hg_config* cfg) {
  if (!c || !cfg) return HG_ERR_BAD_ARG;
  *cfg = c->cfg;
  return HG_OK;
}

hg_status hg_get_progress(hg_ctx* c, hg_progress* out) {
  if (!c || !out) return HG_ERR_BAD_ARG;
  return Guard(c, [&]() {
    c->Sync();  // settles a deferred indicator
    out->time = c->t_fluid;
    out->module_time = c->mod_t;
    out->dt = c->dt;
    out->dt_advection = c->dt_adv;
    out->steps = c->steps;
    out->iteration = c->last_s;
    out->convergence_indicator = (c->cd_iter == 0) ? 1. : c->last_rs;
    out->meshpos[0] = c->meshpos[0];
    out->meshpos[1] = c->meshpos[1];
  });
}

hg_status hg_get_iteration_history(hg_ctx* c, double* rs, int32_t cap, int32_t* n) {
  if (!c || !n || cap < 0 || (cap > 0 && !rs)) return HG_ERR_BAD_ARG;
  return Guard(c, [&]() {
    c->Sync();
    *n = int32_t(c->iter_rs.size());
    for (int32_t k = 0; k < std::min(*n, cap); ++k) rs[k] = c->iter_rs[k];
  });
}

hg_status hg_refresh(hg_ctx* c) {
  if (!c) return HG_ERR_BAD_ARG;
  return Guard(c, [&]() {
    c->fields_checked = false;
    c->flux32_valid = false;
    if (c->open_bc) InitBoundary(c);
    InitialFlux(c);
    UpdateProperties(c);
    InitStats(c);
    c->Sync();
  });
}

hg_status hg_fluid_start_step(hg_ctx* c) {
  if (!c) return HG_ERR_BAD_ARG;
  return Guard(c, [&]() { StartStep(c); });
}

hg_status hg_fluid_iterate(hg_ctx* c, hg_iter_stats* s) {
  if (!c) return HG_ERR_BAD_ARG;
  hg_iter_stats tmp;
  return Guard(c, [&]() { Iterate(c, s ? s : &tmp); });
}

int32_t hg_fluid_is_converged(const hg_ctx* c) { return c && IsConverged(c) ? 1 : 0; }

hg_status hg_fluid_finish_step(hg_ctx* c) {
  if (!c) return HG_ERR_BAD_ARG;
  return Guard(c, [&]() { FinishStep(c); });
}

hg_status hg_advect_substep(hg_ctx* c) {
  if (!c) return HG_ERR_BAD_ARG;
  return Guard(c, [&]() {
    AdvectSubstep(c);
    c->Sync();
  });
}

hg_status hg_update_properties(hg_ctx* c) {
  if (!c) return HG_ERR_BAD_ARG;
  return Guard(c, [&]() {
    UpdateProperties(c);
    AccumulateStats(c);
    c->Sync();
  });
}

hg_status hg_calc_stats(hg_ctx* c, hg_stats* s) {
  if (!c || !s) return HG_ERR_BAD_ARG;
  return Guard(c, [&]() { CalcStats(c, s); });
}

hg_status hg_step(hg_ctx* c, hg_step_stats* s) {
  if (!c) return HG_ERR_BAD_ARG;
  return Guard(c, [&]() {
    hg_step_stats st{};
    HG_CHECK(hipEventRecord(c->ev[0], c->st));
    if (c->cfg.dt_auto) AutoTimeStep(c);
    StartStep(c);
    bool iterated = false, last_deferred = false;
    st.pressure_converged = 1;
    if (c->cfg.fluid_enable) {
      while (!IsConverged(c)) {
        hg_iter_stats is;
        // when the iteration limit ends the loop, the indicator is not
        // needed to decide and is read back at a later synchronisation
        last_deferred = c->fluid_iter + 1 >= c->cfg.num_iterations_limit;
        Iterate(c, &is, last_deferred);
        st.pressure_solve_iters_total += is.pressure_solve_iters;
        if (!last_deferred) {
          st.pressure_solve_diff_max = std::max(st.pressure_solve_diff_max, is.pressure_solve_diff);
          st.pressure_converged = st.pressure_converged && is.pressure_converged;
        }
        iterated = true;
      }
    }
    st.simple_iterations = c->fluid_iter;
    FinishStep(c);
    if (last_deferred) {  // FinishStep's synchronisation settled the last solve
      st.pressure_solve_diff_max =
          std::max(st.pressure_solve_diff_max, c->last_iter.pressure_solve_diff);
      st.pressure_converged = st.pressure_converged && c->last_iter.pressure_converged;
    }
    HG_CHECK(hipEventRecord(c->ev[1], c->st));
    if (c->cfg.advection_enable) {
      while (c->t_adv < c->t_fluid - 0.5 * c->dt_adv) {
        AdvectSubstep(c);
        st.advection_substeps += 1;
      }
    }
    if (c->cfg.heat_enable) HeatStep(c);
    HG_CHECK(hipEventRecord(c->ev[2], c->st));
    UpdateProperties(c);
    AccumulateStats(c);
    HG_CHECK(hipEventRecord(c->ev[3], c->st));
    // no synchronisation at the end of the step: FinishStep's NaN check
    // waited for the SIMPLE loop (settling its indicator), and the work
    // queued since (advection, properties, running stats) overlaps with the
    // host's return and the next call; every readback waits for it, and the
    // region timings are read from the events when asked for
    c->timed = true;
    if (iterated) st.convergence_indicator = c->last_iter.convergence_indicator;
    st.time = c->t_fluid;
    st.dt = c->dt;
    if (s) *s = st;
  });
}

hg_status hg_last_step_timing(const hg_ctx* c, double* f, double* a, double* p) {
  if (!c) return HG_ERR_BAD_ARG;
  float t[3] = {0.f, 0.f, 0.f};
  if (c->timed) {
    if (hipEventSynchronize(c->ev[3]) != hipSuccess) return HG_ERR_HIP;
    for (int k = 0; k < 3; ++k)
      if (hipEventElapsedTime(&t[k], c->ev[k], c->ev[k + 1]) != hipSuccess) return HG_ERR_HIP;
  }
  if (f) *f = t[0];
  if (a) *a = t[1];
  if (p) *p = t[2];
  return HG_OK;
}

hg_status hg_train_pressure_net(hg_ctx* c, const hg_train_opts* opts, hg_train_stats* stats) {
  if (!c || !opts) return HG_ERR_BAD_ARG;
  return Guard(c, [&]() {
    const hg_config& cfg = c->cfg;
    if (cfg.linear_solver_pressure != HG_SOLVER_FCG_LEARNED)
      throw HgError(HG_ERR_BAD_ARG, "linear_solver_pressure is not fcg_learned");
    if (!c->pressure_system)
      throw HgError(HG_ERR_BAD_ARG, "no pressure system yet: run a fluid iteration first");
    if (!hg::LearnedNet::Supported(cfg.nx, cfg.ny))
      throw HgError(HG_ERR_UNSUPPORTED, "learned preconditioner needs Nx, Ny multiples of 8");
    hg::TrainOpts o;
    try {
      o = hg::TrainOptsFromAbi(*opts, c->pinned < 0);
    } catch (const std::invalid_argument& e) {
      throw HgError(HG_ERR_BAD_ARG, e.what());
    }
    if (!c->net) c->net.reset(new hg::LearnedNet(cfg.nx, cfg.ny, c->mg.get()));
    hg::TrainReport r = c->net->Train(*c->cg, o);
    if (stats) {
      stats->pairs = r.pairs;
      stats->steps = r.steps;
      stats->loss_first = r.loss_first;
      stats->loss_last = r.loss_last;
      stats->ms = r.ms;
      stats->accepted = r.accepted ? 1 : 0;
      stats->val_iters = r.val_iters;
      stats->val_iters_baseline = r.val_iters_baseline;
      stats->val_ms = r.val_ms;
      stats->val_ms_baseline = r.val_ms_baseline;
    }
  });
}

hg_status hg_batch_run(hg_ctx* const* ctxs, int32_t n, int32_t steps, int32_t threads,
                       hg_step_stats* stats, double* wall_ms) {
  if (!ctxs || n < 0 || steps < 0 || threads < 0) return HG_ERR_BAD_ARG;
  for (int32_t i = 0; i < n; ++i)
    if (!ctxs[i]) return HG_ERR_BAD_ARG;
  {
    // a context listed twice would be stepped by two threads at once
    std::vector<hg_ctx*> sorted(ctxs, ctxs + n);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end()) return HG_ERR_BAD_ARG;
  }
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<hg_status> rc(size_t(n), HG_OK);
  std::atomic<int32_t> next{0};
  auto worker = [&]() {
    for (int32_t i = next.fetch_add(1); i < n; i = next.fetch_add(1)) {
      hg_step_stats last{};
      for (int32_t s = 0; s < steps && rc[i] == HG_OK; ++s) rc[i] = hg_step(ctxs[i], &last);
      // hg_step returns with its last kernels queued: the case is done when
      // its stream is
      if (rc[i] == HG_OK) rc[i] = Guard(ctxs[i], [&]() { ctxs[i]->Sync(); });
      if (stats) stats[i] = last;
    }
  };
  const int32_t nt = std::max<int32_t>(1, std::min<int32_t>(threads > 0 ? threads : n, n));
  std::vector<std::thread> pool;
  for (int32_t t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  if (wall_ms)
    *wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                   .count();
  for (int32_t i = 0; i < n; ++i)
    if (rc[i] != HG_OK) return rc[i];
  return HG_OK;
}

hg_status hg_time_kernel(hg_ctx* c, const char* kernel, int32_t reps, double* avg_ms) {
  if (!c || !kernel || reps < 1 || !avg_ms) return HG_ERR_BAD_ARG;
  return Guard(c, [&]() {
    const std::string k(kernel);
    std::function<void()> launch;
    if (k == "KAdvStep") {
      // reads the phase fields and the flux, writes the scratch layer pd_new
      const double* u[3] = {c->pd[0], c->np > 1 ? c->pd[1] : nullptr,
                            c->np > 2 ? c->pd[2] : nullptr};
      double* un[3] = {c->pd_new[0], c->np > 1 ? c->pd_new[1] : nullptr,
                       c->np > 2 ? c->pd_new[2] : nullptr};
      if (c->mixed) {
        if (!c->flux32_valid) {
          hg::LaunchToFloat(c->st, c->nf, c->fl_ic, c->f32.flux);
          c->flux32_valid = true;
        }
        launch = [=]() {
          hg::LaunchAdvStep(c->st, c->geo, c->np, u, c->f32.flux, c->dt_adv, un);
        };
      } else {
        launch = [=]() { hg::LaunchAdvStep(c->st, c->geo, c->np, u, c->fl_ic, c->dt_adv, un); };
      }
    } else if (k == "KMomentum" && c->mixed) {
      hg::MomentumArgsF m{};
      auto& f = c->f32;
      m.u0 = c->u_ip[0];
      m.u1 = c->u_ip[1];
      m.g0x = f.g0x; m.g0y = f.g0y; m.g1x = f.g1x; m.g1y = f.g1y;
      m.gpx = f.gpx; m.gpy = f.gpy; m.frx = f.frx; m.fry = f.fry;
      m.ffmu = f.ffmu;
      m.flux = c->fl_ip;
      m.rho = c->rho_s;
      m.utp0 = c->u_tp[0]; m.utp1 = c->u_tp[1]; m.utc0 = c->u_tc[0]; m.utc1 = c->u_tc[1];
      m.bdf0 = c->bdf[0]; m.bdf1 = c->bdf[1]; m.bdf2 = c->bdf[2];
      m.relax = c->cfg.velocity_relaxation_factor;
      m.s = f.s; m.w = f.w; m.c = f.c; m.e = f.e; m.n = f.n; m.k0 = f.k0; m.k1 = f.k1;
      m.ap = c->ap;
      launch = [=]() { hg::LaunchMomentum(c->st, c->geo, c->bc, m); };
    } else if (k == "KMomentum") {
      // the rows of the last SIMPLE iteration's inputs, rewritten in place
      hg::MomentumArgs m{};
      m.u0 = c->u_ip[0];
      m.u1 = c->u_ip[1];
      m.g0x = c->g0x; m.g0y = c->g0y; m.g1x = c->g1x; m.g1y = c->g1y;
      m.gpx = c->gpx; m.gpy = c->gpy; m.frx = c->frx; m.fry = c->fry;
      m.ffmu = c->ffmu;
      m.flux = c->fl_ip;
      m.rho = c->rho_s;
      m.utp0 = c->u_tp[0]; m.utp1 = c->u_tp[1]; m.utc0 = c->u_tc[0]; m.utc1 = c->u_tc[1];
      m.bdf0 = c->bdf[0]; m.bdf1 = c->bdf[1]; m.bdf2 = c->bdf[2];
      m.relax = c->cfg.velocity_relaxation_factor;
      m.s = c->As; m.w = c->Aw; m.c = c->Ac; m.e = c->Ae; m.n = c->An; m.k0 = c->k0;
      m.k1 = c->k1; m.ap = c->ap;
      launch = [=]() { hg::LaunchMomentum(c->st, c->geo, c->bc, m); };
    } else {
      throw HgError(HG_ERR_BAD_ARG, "hg_time_kernel: unknown kernel " + k);
    }
    launch();
    HG_CHECK(hipEventRecord(c->kev[0], c->st));
    for (int32_t i = 0; i < reps; ++i) launch();
    HG_CHECK(hipEventRecord(c->kev[1], c->st));
    HG_CHECK(hipEventSynchronize(c->kev[1]));
    float ms = 0.f;
    HG_CHECK(hipEventElapsedTime(&ms, c->kev[0], c->kev[1]));
    *avg_ms = double(ms) / reps;
  });
}

}  // extern "C"
