// noalloc.hip -- hg_stencil_solve_noalloc: the reference's row solvers on
// caller-owned device arrays, with no device allocation of any kind.
//
// hg_stencil_solve copies host rows into device buffers it allocates per
// call.  Callers that cannot allocate (code under HIP graph capture, tight
// serving loops) pass device arrays instead, and every work array the
// solvers need -- the second Jacobi iterate, the red-black sweep's black
// values, lu_relaxed's correction and evaluated rows, the iteration info --
// lives in one file-scope device array, `scratch`, of kScratchBytes.  Each
// GPU holds its own copy of that array until the process ends; its address
// on the current device is looked up once per device with
// hipGetSymbolAddress and kept.  Calls on one device share the array, so
// calls that overlap on one device race; every array that must start at
// zero is cleared by a fill kernel at the start of each call.

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/hydro_gpu.h"
#include "../../include/hydro_noalloc.h"
#include "kernels.h"

namespace {

constexpr size_t kScratchBytes = size_t(HG_NOALLOC_SCRATCH_BYTES);
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each device (the array's name is a host
// stand-in, not an address); looked up once per device.
void* g_scratch_ptr[kMaxDevices] = {};
std::once_flag g_scratch_once[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices] = {};

hg_status Fail(hg_status rc, const std::string& msg) {
  hg::SetThreadError("hg_stencil_solve_noalloc: " + msg);
  std::fprintf(stderr, "hg_stencil_solve_noalloc: %s\n", msg.c_str());
  return rc;
}

template <class T>
__global__ void KFillZero(T* __restrict__ p, long n) {
  for (long i = blockIdx.x * long(blockDim.x) + threadIdx.x; i < n;
       i += long(gridDim.x) * blockDim.x)
    p[i] = T(0);
}

template <class T>
void Zero(hipStream_t st, T* p, long n) {
  const int blocks = int(std::min<long>((n + 255) / 256, 1024));
  KFillZero<T><<<std::max(blocks, 1), 256, 0, st>>>(p, n);
}

inline size_t Up256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

extern "C" {

hg_status hg_stencil_solve_noalloc(int32_t nx, int32_t ny, const double* s, const double* w,
                                   const double* c, const double* e, const double* n,
                                   const double* k, double* x, const hg_solve_opts* opts,
                                   hg_solve_stats* stats, void* stream) {
  if (nx < 1 || ny < 1 || !s || !w || !c || !e || !n || !k || !x || !opts)
    return Fail(HG_ERR_BAD_ARG, "null array or empty grid");
  const int32_t solver = opts->solver;
  const bool iterative = solver == HG_SOLVER_GAUSS_SEIDEL || solver == HG_SOLVER_JACOBI ||
                         solver == HG_SOLVER_LU_RELAXED;
  if (!(iterative || solver == HG_SOLVER_LU || solver == HG_SOLVER_LU_RB))
    return Fail(HG_ERR_UNSUPPORTED,
                "solver must be lu, lu_rb, gauss_seidel, jacobi or lu_relaxed");
  if (nx > 4096) return Fail(HG_ERR_BAD_ARG, "nx above 4096 (one workgroup's columns)");
  const size_t N = size_t(nx) * size_t(ny);
  const int limit = iterative ? opts->max_iters : 0;
  if (limit < 0) return Fail(HG_ERR_BAD_ARG, "max_iters < 0");
  // layout: [info: 2 x u64 | jacobi state: 4 x int] [jacobi diffs: limit + 2]
  //         [four cell arrays of N doubles]
  const size_t off_diff = 256;
  const size_t off_arr = off_diff + Up256(size_t(limit + 2) * 8);
  const size_t arr = Up256(N * 8);
  const size_t need = off_arr + 4 * arr;
  if (need > kScratchBytes)
    return Fail(HG_ERR_BAD_ARG, "grid of " + std::to_string(N) + " cells with max_iters " +
                                    std::to_string(limit) + " needs " + std::to_string(need) +
                                    " scratch bytes; the maximum is " +
                                    std::to_string(kScratchBytes));
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return Fail(HG_ERR_HIP, "no current HIP device");
  std::call_once(g_scratch_once[dev], [dev]() {
    void* p = nullptr;
    g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_ptr[dev] = (g_scratch_err[dev] == hipSuccess) ? p : nullptr;
  });
  if (!g_scratch_ptr[dev])
    return Fail(HG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                hipGetErrorString(g_scratch_err[dev]));
  unsigned char* base = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  unsigned long long* info = reinterpret_cast<unsigned long long*>(base);
  int* jstate = reinterpret_cast<int*>(base + 16);
  unsigned long long* jdiff = reinterpret_cast<unsigned long long*>(base + off_diff);
  double* a0 = reinterpret_cast<double*>(base + off_arr);
  double* a1 = reinterpret_cast<double*>(base + off_arr + arr);
  double* a2 = reinterpret_cast<double*>(base + off_arr + 2 * arr);
  double* a3 = reinterpret_cast<double*>(base + off_arr + 3 * arr);

  hipStream_t st = static_cast<hipStream_t>(stream);
  // the last-error record may hold an earlier call's error on this thread:
  // clear it, so the check after the launches sees only this call's
  (void)hipGetLastError();
  hg::Rows rows{s, w, c, e, n};
  int iters = 0;
  double diff = 0.;
  auto check = [](hipError_t err) { return err == hipSuccess; };
  switch (solver) {
    case HG_SOLVER_LU:
      hg::LaunchLuWavefront(st, nx, ny, rows, k, nullptr, x, nullptr, nullptr, nullptr, nullptr,
                            nullptr);
      break;
    case HG_SOLVER_LU_RB:
      // one right-hand side: the second lane solves it again into a3; the
      // delta-form base is a zeroed array
      Zero(st, a0, long(N));
      hg::LaunchLuRb(st, nx, ny, rows, k, k, a1, a2, a0, a0, x, a3);
      break;
    case HG_SOLVER_GAUSS_SEIDEL:
      hg::LaunchGsWavefront(st, nx, ny, rows, k, x, opts->tolerance, limit,
                            opts->relaxation_factor, info);
      break;
    case HG_SOLVER_LU_RELAXED:
      hg::LaunchLuRelaxed(st, nx, ny, rows, k, x, a0, a1, opts->tolerance, limit,
                          opts->relaxation_factor, info);
      break;
    case HG_SOLVER_JACOBI: {
      hg::JacobiWork jw{a0, a1, jdiff, jstate};
      Zero(st, a0, long(N));
      Zero(st, jdiff, long(limit + 2));
      Zero(st, jstate, 4L);
      int hstate[4] = {0, 0, 0, 0};
      int first = 0;
      while (first <= limit) {
        const int cnt = std::min(64, limit + 1 - first);
        hg::LaunchJacobiSweeps(st, nx, ny, rows, k, jw, first, cnt, opts->tolerance, limit,
                               opts->relaxation_factor);
        first += cnt;
        if (!check(hipMemcpyAsync(hstate, jstate, sizeof(hstate), hipMemcpyDeviceToHost, st)) ||
            !check(hipStreamSynchronize(st)))
          return Fail(HG_ERR_HIP, "jacobi state readback failed");
        if (hstate[0]) break;
      }
      iters = hstate[2];
      const int last = std::max(0, std::min(iters, std::min(first, limit + 1) - 1));
      unsigned long long bits = 0;
      if (!check(hipMemcpyAsync(x, hstate[3] ? a1 : a0, N * 8, hipMemcpyDeviceToDevice, st)) ||
          !check(hipMemcpyAsync(&bits, jdiff + last, 8, hipMemcpyDeviceToHost, st)) ||
          !check(hipStreamSynchronize(st)))
        return Fail(HG_ERR_HIP, "jacobi result copy failed");
      std::memcpy(&diff, &bits, 8);
      break;
    }
    default:
      break;
  }
  if (solver == HG_SOLVER_GAUSS_SEIDEL || solver == HG_SOLVER_LU_RELAXED) {
    unsigned long long h[2] = {0, 0};
    if (!check(hipMemcpyAsync(h, info, 16, hipMemcpyDeviceToHost, st)) ||
        !check(hipStreamSynchronize(st)))
      return Fail(HG_ERR_HIP, "iteration info readback failed");
    iters = int(h[0]);
    std::memcpy(&diff, &h[1], 8);
  } else if (!check(hipStreamSynchronize(st))) {
    return Fail(HG_ERR_HIP, "stream synchronisation failed");
  }
  if (!check(hipGetLastError())) return Fail(HG_ERR_HIP, "kernel launch failed");
  if (stats) {
    stats->iterations = iters;
    stats->final_diff = diff;
    stats->converged = iterative ? !(diff > opts->tolerance) : 1;
    stats->rel_residual = -1.;  // not computed: the rows stay on the device
    stats->ms = 0.;
  }
  return HG_OK;
}

}  // extern "C"
