// poisson.hip -- solve-level entry points of the C-ABI.
//
// hg_stencil_solve replaces LinearSolver<Scal,IdxCell,Expr>::Solve
// (linear.hpp:302-310) for the reference's own solvers on 5-point rows;
// hg_poisson_* keep a variable-coefficient pressure operator resident in
// HBM for repeated CG / FCG solves.

#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hydro_gpu.h"
#include "cg.h"
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

struct DevBuf {
  double* p = nullptr;
  explicit DevBuf(size_t n) { HG_CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * 8)); }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

struct hg_poisson {
  int nx = 0, ny = 0;
  hipStream_t st = nullptr;
  std::unique_ptr<hg::CgSolver> cg;
  std::unique_ptr<hg::LearnedNet> net;
  std::unique_ptr<hg::MgPrecond> mg;
  std::unique_ptr<hg::MgPrecond64> mg64;  // built on the first FP64 fcg_mg solve
  double *b = nullptr, *x = nullptr;
  ~hg_poisson() {
    if (st) (void)hipStreamSynchronize(st);
    cg.reset();
    net.reset();
    mg.reset();
    mg64.reset();
    if (b) (void)hipFree(b);
    if (x) (void)hipFree(x);
    if (st) (void)hipStreamDestroy(st);
  }
};

extern "C" {

hg_status hg_stencil_solve(int32_t nx, int32_t ny, const double* s, const double* w,
                           const double* c, const double* e, const double* n, const double* k,
                           double* x, const hg_solve_opts* opts, hg_solve_stats* stats) {
  if (nx < 1 || ny < 1 || !s || !w || !c || !e || !n || !k || !x || !opts)
    return HG_ERR_BAD_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return HG_ERR_HIP;
  try {
    const size_t N = size_t(nx) * ny;
    hipStream_t st;
    HG_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    std::unique_ptr<void, std::function<void(void*)>> guard(
        (void*)st, [](void* p) { (void)hipStreamDestroy((hipStream_t)p); });
    DevBuf ds(N), dw(N), dc(N), de(N), dn(N), dk(N), dx(N), dx2(N), dz(N), dout(N), dtmp(N);
    DevBuf info(2), jdiff(size_t(std::max(opts->max_iters, 0)) + 2), jstate(2);
    const double* src[6] = {s, w, c, e, n, k};
    double* dst[6] = {ds.p, dw.p, dc.p, de.p, dn.p, dk.p};
    for (int q = 0; q < 6; ++q)
      HG_CHECK(hipMemcpyAsync(dst[q], src[q], N * 8, hipMemcpyHostToDevice, st));
    HG_CHECK(hipMemsetAsync(dz.p, 0, N * 8, st));
    hg::Rows rows{ds.p, dw.p, dc.p, de.p, dn.p};
    hipEvent_t e0, e1;
    HG_CHECK(hipEventCreate(&e0));
    HG_CHECK(hipEventCreate(&e1));
    HG_CHECK(hipEventRecord(e0, st));
    int iters = 0;
    double diff = 0.;
    const double* result = dx.p;
    switch (opts->solver) {
      case HG_SOLVER_LU:
        hg::LaunchLuWavefront(st, nx, ny, rows, dk.p, nullptr, dx.p, nullptr, nullptr, nullptr,
                              nullptr, nullptr);
        break;
      case HG_SOLVER_LU_RB:
        hg::LaunchLuRb(st, nx, ny, rows, dk.p, dk.p, dx.p, dx2.p, dz.p, dz.p, dout.p, dtmp.p);
        result = dout.p;
        break;
      case HG_SOLVER_GAUSS_SEIDEL: {
        hg::LaunchGsWavefront(st, nx, ny, rows, dk.p, dx.p, opts->tolerance, opts->max_iters,
                              opts->relaxation_factor, (unsigned long long*)info.p);
        unsigned long long h[2];
        HG_CHECK(hipMemcpyAsync(h, info.p, 16, hipMemcpyDeviceToHost, st));
        HG_CHECK(hipStreamSynchronize(st));
        iters = int(h[0]);
        std::memcpy(&diff, &h[1], 8);
        break;
      }
      case HG_SOLVER_LU_RELAXED: {
        hg::LaunchLuRelaxed(st, nx, ny, rows, dk.p, dx.p, dx2.p, dtmp.p, opts->tolerance,
                            opts->max_iters, opts->relaxation_factor, (unsigned long long*)info.p);
        unsigned long long h[2];
        HG_CHECK(hipMemcpyAsync(h, info.p, 16, hipMemcpyDeviceToHost, st));
        HG_CHECK(hipStreamSynchronize(st));
        iters = int(h[0]);
        std::memcpy(&diff, &h[1], 8);
        break;
      }
      case HG_SOLVER_JACOBI: {
        hg::JacobiWork jw{dx.p, dx2.p, (unsigned long long*)jdiff.p, (int*)jstate.p};
        HG_CHECK(hipMemsetAsync(dx.p, 0, N * 8, st));
        HG_CHECK(hipMemsetAsync(jdiff.p, 0, (size_t(opts->max_iters) + 2) * 8, st));
        HG_CHECK(hipMemsetAsync(jstate.p, 0, 16, st));
        int first = 0;
        int hstate[4] = {0, 0, 0, 0};
        while (first <= opts->max_iters) {
          int cnt = std::min(64, opts->max_iters + 1 - first);
          hg::LaunchJacobiSweeps(st, nx, ny, rows, dk.p, jw, first, cnt, opts->tolerance,
                                 opts->max_iters, opts->relaxation_factor);
          first += cnt;
          HG_CHECK(hipMemcpyAsync(hstate, jstate.p, 16, hipMemcpyDeviceToHost, st));
          HG_CHECK(hipStreamSynchronize(st));
          if (hstate[0]) break;
        }
        iters = hstate[2];
        result = hstate[3] ? dx2.p : dx.p;
        unsigned long long bits = 0;
        int last = std::max(0, std::min(iters, std::min(first, opts->max_iters + 1) - 1));
        HG_CHECK(hipMemcpyAsync(&bits, (unsigned long long*)jdiff.p + last, 8,
                                hipMemcpyDeviceToHost, st));
        HG_CHECK(hipStreamSynchronize(st));
        std::memcpy(&diff, &bits, 8);
        break;
      }
      default:
        return HG_ERR_UNSUPPORTED;
    }
    HG_CHECK(hipEventRecord(e1, st));
    HG_CHECK(hipMemcpyAsync(x, result, N * 8, hipMemcpyDeviceToHost, st));
    HG_CHECK(hipStreamSynchronize(st));
    float ms = 0.f;
    HG_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (stats) {
      stats->iterations = iters;
      stats->final_diff = diff;
      stats->converged = !(diff > opts->tolerance);
      stats->ms = ms;
      // residual of the returned solution, FP64 on the host (diagnostic)
      double rn = 0., bn = 0.;
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          size_t r = size_t(i) + size_t(nx) * j;
          double v = k[r] + c[r] * x[r];
          if (j > 0) v += s[r] * x[r - nx];
          if (i > 0) v += w[r] * x[r - 1];
          if (i < nx - 1) v += e[r] * x[r + 1];
          if (j < ny - 1) v += n[r] * x[r + nx];
          rn += v * v;
          bn += k[r] * k[r];
        }
      stats->rel_residual = bn > 0 ? std::sqrt(rn / bn) : 0.;
    }
    return HG_OK;
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "hg_stencil_solve: %s\n", ex.what());
    return HG_ERR_HIP;
  }
}

hg_status hg_poisson_create(int32_t nx, int32_t ny, const double* kx, const double* ky,
                            hg_poisson** out) {
  if (nx < 2 || ny < 2 || !kx || !ky || !out) return HG_ERR_BAD_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return HG_ERR_HIP;
  try {
    std::unique_ptr<hg_poisson> ps(new hg_poisson());
    ps->nx = nx;
    ps->ny = ny;
    HG_CHECK(hipStreamCreateWithFlags(&ps->st, hipStreamNonBlocking));
    ps->cg.reset(new hg::CgSolver(nx, ny, ps->st));
    const size_t nfx = size_t(nx + 1) * ny, nfy = size_t(nx) * (ny + 1), N = size_t(nx) * ny;
    DevBuf dkx(nfx), dky(nfy);
    HG_CHECK(hipMemcpyAsync(dkx.p, kx, nfx * 8, hipMemcpyHostToDevice, ps->st));
    HG_CHECK(hipMemcpyAsync(dky.p, ky, nfy * 8, hipMemcpyHostToDevice, ps->st));
    ps->cg->SetOperator(dkx.p, dky.p);
    ps->mg.reset(new hg::MgPrecond(nx, ny, ps->cg->kx32(), ps->cg->ky32(), ps->cg->dinv32()));
    ps->mg->Build(ps->st);
    HG_CHECK(hipMalloc(&ps->b, N * 8));
    HG_CHECK(hipMalloc(&ps->x, N * 8));
    HG_CHECK(hipStreamSynchronize(ps->st));
    *out = ps.release();
    return HG_OK;
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "hg_poisson_create: %s\n", ex.what());
    return HG_ERR_HIP;
  }
}

void hg_poisson_destroy(hg_poisson* ps) { delete ps; }

hg_status hg_poisson_solve(hg_poisson* ps, const double* rhs, double* p,
                           const hg_solve_opts* opts, hg_solve_stats* stats) {
  if (!ps || !rhs || !p || !opts) return HG_ERR_BAD_ARG;
  if (opts->solver != HG_SOLVER_CG && opts->solver != HG_SOLVER_FCG_LEARNED &&
      opts->solver != HG_SOLVER_FCG_MG)
    return HG_ERR_UNSUPPORTED;
  try {
    const size_t N = size_t(ps->nx) * ps->ny;
    HG_CHECK(hipMemcpyAsync(ps->b, rhs, N * 8, hipMemcpyHostToDevice, ps->st));
    hg::Precond* pc = nullptr;
    if (opts->solver == HG_SOLVER_FCG_LEARNED)
      pc = (ps->net && ps->net->Trained()) ? static_cast<hg::Precond*>(ps->net.get())
                                           : static_cast<hg::Precond*>(ps->mg.get());
    else if (opts->solver == HG_SOLVER_FCG_MG && opts->precision == HG_PREC_FP64) {
      // the same V-cycle with FP64 levels (what FP32 buys, measured)
      if (!ps->mg64) {
        ps->mg64.reset(new hg::MgPrecond64(ps->nx, ps->ny, ps->cg->kx64(), ps->cg->ky64(),
                                           ps->cg->dinv64()));
        ps->mg64->Build(ps->st);
      }
      pc = ps->mg64.get();
    } else if (opts->solver == HG_SOLVER_FCG_MG)
      pc = ps->mg.get();
    hg::CgResult r = ps->cg->Solve(ps->b, ps->x, opts->tolerance, opts->max_iters,
                                   opts->precision, true, pc);
    HG_CHECK(hipMemcpyAsync(p, ps->x, N * 8, hipMemcpyDeviceToHost, ps->st));
    HG_CHECK(hipStreamSynchronize(ps->st));
    if (stats) {
      stats->iterations = r.iterations;
      stats->final_diff = r.rel_residual;
      stats->rel_residual = r.rel_residual;
      stats->converged = r.converged ? 1 : 0;
      stats->ms = r.ms;
    }
    return r.converged ? HG_OK : HG_ERR_NOT_CONVERGED;
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "hg_poisson_solve: %s\n", ex.what());
    return HG_ERR_HIP;
  }
}

hg_status hg_poisson_time_kernel(hg_poisson* ps, const char* kernel, int32_t reps,
                                 double* avg_ms) {
  if (!ps || !kernel || reps < 1 || !avg_ms) return HG_ERR_BAD_ARG;
  try {
    *avg_ms = ps->mg->TimeKernel(kernel, reps, ps->b, ps->st);
    return HG_OK;
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "hg_poisson_time_kernel: %s\n", ex.what());
    return HG_ERR_BAD_ARG;
  }
}

void hg_train_opts_default(hg_train_opts* o) {
  if (!o) return;
  hg::TrainOpts d;
  o->rhs_count = d.rhs_count;
  o->mg_iters = d.mg_iters;
  o->steps = d.steps;
  o->lr = d.lr;
  o->seed = d.seed;
  o->flags = (d.linear ? HG_TRAIN_LINEAR : 0) | (d.standalone ? HG_TRAIN_STANDALONE : 0) |
             (d.tied ? HG_TRAIN_TIED : 0);
}

hg_status hg_poisson_train(hg_poisson* ps, const hg_train_opts* opts, hg_train_stats* stats) {
  if (!ps || !opts) return HG_ERR_BAD_ARG;
  if (!hg::LearnedNet::Supported(ps->nx, ps->ny)) return HG_ERR_UNSUPPORTED;
  try {
    hg::TrainOpts o = hg::TrainOptsFromAbi(*opts, true);
    if (!ps->net) ps->net.reset(new hg::LearnedNet(ps->nx, ps->ny, ps->mg.get()));
    hg::TrainReport r = ps->net->Train(*ps->cg, o);
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
    return HG_OK;
  } catch (const std::invalid_argument& ex) {
    std::fprintf(stderr, "hg_poisson_train: %s\n", ex.what());
    return HG_ERR_BAD_ARG;
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "hg_poisson_train: %s\n", ex.what());
    return HG_ERR_HIP;
  }
}

}  // extern "C"
