// solvers.hip -- 5-point linear solvers of the reference, on the GPU.
//
// Rows: s x[r-nx] + w x[r-1] + c x[r] + e x[r+1] + n x[r+nx] + k = 0, raw
// index r = i + nx*j.  The reference's "lu" and "gauss_seidel" sweep the
// cells in raw order (linear.hpp:536-563, 692-709), so cell (i,j) needs the
// new values of (i-1,j) and (i,j-1): the anti-diagonal i+j is a wavefront.
// One workgroup runs the whole solve: thread t owns a block of columns and
// handles row j = step - t at each step; the west value crosses threads
// through LDS, the south value stays in registers, and one barrier separates
// steps.  Each cell computes exactly the reference's expression, so the
// result equals the serial sweep bit for bit.  Compiled with
// -ffp-contract=off.

#include <type_traits>

#include "kernels.h"

namespace hg {

namespace {

constexpr int kMaxCw = 4;        // columns per thread (nx <= 4096)
constexpr int kMaxThreads = 1024;

inline int ThreadsFor(int nx, int* cw) {
  int c = (nx + kMaxThreads - 1) / kMaxThreads;
  if (c < 1) c = 1;
  *cw = c;
  int t = (nx + c - 1) / c;
  return ((t + 63) / 64) * 64;
}

// Barrier between wavefront steps that orders only the LDS exchange: the
// global loads a step has in flight (the next row's operands) are not waited
// for, since no thread writes what they read before that row's step.
__device__ __forceinline__ void LdsBarrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Operands of one row of the LU sweep for a thread's CW columns and NR
// right-hand sides: forward, the south / west coefficients, the diagonal and
// the constants; backward, the north / east coefficients, the diagonal and
// the forward pass's values.  Loaded one step ahead (nothing a row's step
// reads is written by another thread in the same pass; the backward pass
// reads the thread's own forward results); with one column per thread every
// load is issued at a clamped index, as in the SOR kernel below.
template <int NR, int CW>
struct LuRow {
  double a[CW], b[CW], c[CW], v[NR][CW];
};

template <int NR, int CW, bool FWD>
__device__ __forceinline__ void LuLoad(LuRow<NR, CW>& o, int nx, int ny, int c0, int j,
                                       const Rows& a, const double* const (&kk)[2],
                                       double* const (&xx)[2]) {
  if constexpr (CW == 1) {
    const int jc = min(max(j, 0), ny - 1), i = min(c0, nx - 1), r = i + nx * jc;
    o.a[0] = FWD ? a.s[r] : a.n[r];
    o.b[0] = FWD ? a.w[r] : a.e[r];
    o.c[0] = a.c[r];
#pragma unroll
    for (int q = 0; q < NR; ++q) o.v[q][0] = FWD ? kk[q][r] : xx[q][r];
  } else {
    if (j < 0 || j >= ny) return;
#pragma unroll
    for (int m = 0; m < CW; ++m) {
      const int i = c0 + m;
      if (i >= nx) break;
      const int r = i + nx * j;
      o.a[m] = FWD ? a.s[r] : a.n[r];
      o.b[m] = FWD ? a.w[r] : a.e[r];
      o.c[m] = a.c[r];
#pragma unroll
      for (int q = 0; q < NR; ++q) o.v[q][m] = FWD ? kk[q][r] : xx[q][r];
    }
  }
}

// Forward half of the symmetric sweep: x = (-k - s x_S - w x_W) / c.
// Backward half: x -= (n x_N + e x_E) / c.  NR = number of right-hand sides.
template <int NR, int CW>
__global__ void __launch_bounds__(kMaxThreads)
KLuWavefront(int nx, int ny, Rows a, const double* __restrict__ k0,
             const double* __restrict__ k1, double* __restrict__ x0, double* __restrict__ x1,
             const double* __restrict__ base0, const double* __restrict__ base1,
             double* __restrict__ out0, double* __restrict__ out1) {
  extern __shared__ double lds[];  // [2 parities][NR][T]
  const int T = blockDim.x;
  const int t = threadIdx.x;
  const int c0 = t * CW;
  const int nc = max(0, min(nx, c0 + CW) - c0);
  const double* const kk[2] = {k0, k1};
  double* const xx[2] = {x0, x1};
  const int nsteps = ny + T - 1;
  LuRow<NR, CW> ra, rb;
  // ---- forward: row j = step - t, columns left to right
  double south[NR][CW];
  auto Fwd = [&](int step, const LuRow<NR, CW>& o) {
    const int j = step - t;
    const int par = step & 1;
    if (nc > 0 && j >= 0 && j < ny) {
      double west[NR];
#pragma unroll
      for (int q = 0; q < NR; ++q) west[q] = (t > 0) ? lds[((par ^ 1) * NR + q) * T + t - 1] : 0.;
#pragma unroll
      for (int m = 0; m < CW; ++m) {
        if (m >= nc) break;
        const int i = c0 + m, r = i + nx * j;
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          double sum = 0;
          if (j > 0) sum += o.a[m] * south[q][m];
          if (i > 0) sum += o.b[m] * west[q];
          const double v = (-o.v[q][m] - sum) / o.c[m];
          xx[q][r] = v;
          south[q][m] = v;
          west[q] = v;
        }
      }
#pragma unroll
      for (int q = 0; q < NR; ++q) lds[(par * NR + q) * T + t] = west[q];
    }
    LdsBarrier();
  };
  LuLoad<NR, CW, true>(ra, nx, ny, c0, -t, a, kk, xx);
  for (int step = 0; step < nsteps; step += 2) {
    LuLoad<NR, CW, true>(rb, nx, ny, c0, step + 1 - t, a, kk, xx);
    Fwd(step, ra);
    if (step + 1 < nsteps) {
      LuLoad<NR, CW, true>(ra, nx, ny, c0, step + 2 - t, a, kk, xx);
      Fwd(step + 1, rb);
    }
  }
  // ---- backward: thread T-1 first, rows from the top, columns right to left
  double north[NR][CW];
  const int lag = T - 1 - t;
  auto Bwd = [&](int step, const LuRow<NR, CW>& o) {
    const int j = ny - 1 - (step - lag);
    const int par = step & 1;
    if (nc > 0 && j >= 0 && j < ny) {
      double east[NR];
#pragma unroll
      for (int q = 0; q < NR; ++q)
        east[q] = (t < T - 1) ? lds[((par ^ 1) * NR + q) * T + t + 1] : 0.;
#pragma unroll
      for (int m = CW - 1; m >= 0; --m) {
        if (m >= nc) continue;
        const int i = c0 + m, r = i + nx * j;
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          double sum = 0;
          if (j < ny - 1) sum += o.a[m] * north[q][m];
          if (i < nx - 1) sum += o.b[m] * east[q];
          double v = o.v[q][m];
          v -= sum / o.c[m];
          xx[q][r] = v;
          north[q][m] = v;
          east[q] = v;
        }
      }
#pragma unroll
      for (int q = 0; q < NR; ++q) lds[(par * NR + q) * T + t] = east[q];
    }
    LdsBarrier();
  };
  LuLoad<NR, CW, false>(ra, nx, ny, c0, ny - 1 + lag, a, kk, xx);
  for (int step = 0; step < nsteps; step += 2) {
    LuLoad<NR, CW, false>(rb, nx, ny, c0, ny - 1 - (step + 1 - lag), a, kk, xx);
    Bwd(step, ra);
    if (step + 1 < nsteps) {
      LuLoad<NR, CW, false>(ra, nx, ny, c0, ny - 1 - (step + 2 - lag), a, kk, xx);
      Bwd(step + 1, rb);
    }
  }
  __syncthreads();  // every thread's x before the delta-form update reads it
  // ---- delta-form update out = base + x
  if (base0) {
    const double* bb[2] = {base0, base1};
    double* oo[2] = {out0, out1};
    for (int r = t; r < nx * ny; r += T)
      for (int q = 0; q < NR; ++q) oo[q][r] = bb[q][r] + xx[q][r];
  }
}

// One cell row's operands for a thread's CW columns: the row coefficients,
// the constant, and the old values of the cell and of its east and north
// neighbours.  Within a sweep those three are still the previous sweep's
// when the row's step comes: the east cell belongs to this thread (later in
// the row) or to the next thread (one step later), the north cell to the
// next row of this thread.
template <int CW>
struct GsRow {
  double s[CW], w[CW], e[CW], n[CW], c[CW], k[CW], xo[CW], xe[CW], xn[CW];
};

// With one column per thread every load is issued, with the row and column
// clamped into the grid (a clamped value is never used: the sweep skips rows
// outside the grid and columns past nx, and tests i < nx - 1, j < ny - 1
// before using the east and north values), so the number of loads in flight
// is the same on every path and the wait for a row's operands leaves the next
// row's loads in flight (191 against 220 us per 256 x 128 sweep).  Wider
// threads load only what they use: the unconditional form needs more
// registers than they have and spills (2062 against 592 us, 2048 x 8).
template <int CW>
__device__ __forceinline__ void GsLoad(GsRow<CW>& o, int nx, int ny, int c0, int j,
                                       const Rows& a, const double* __restrict__ k,
                                       const double* __restrict__ x) {
  if constexpr (CW == 1) {
    const int jc = min(max(j, 0), ny - 1);
    const int i = min(c0, nx - 1), r = i + nx * jc;
    o.s[0] = a.s[r];
    o.w[0] = a.w[r];
    o.e[0] = a.e[r];
    o.n[0] = a.n[r];
    o.c[0] = a.c[r];
    o.k[0] = k[r];
    o.xo[0] = x[r];
    o.xe[0] = x[i < nx - 1 ? r + 1 : r];
    o.xn[0] = x[jc < ny - 1 ? r + nx : r];
  } else {
    if (j < 0 || j >= ny) return;
#pragma unroll
    for (int m = 0; m < CW; ++m) {
      const int i = c0 + m;
      if (i >= nx) break;
      const int r = i + nx * j;
      o.s[m] = a.s[r];
      o.w[m] = a.w[r];
      o.e[m] = a.e[r];
      o.n[m] = a.n[r];
      o.c[m] = a.c[r];
      o.k[m] = k[r];
      o.xo[m] = x[r];
      o.xe[m] = i < nx - 1 ? x[r + 1] : 0.;
      o.xn[m] = j < ny - 1 ? x[r + nx] : 0.;
    }
  }
}

// Lexicographic SOR, every sweep in this workgroup (linear.hpp:685-715).
// Thread t owns columns c0 .. c0 + CW - 1 and takes row j = step - t at each
// step; the row of the next step is loaded while this one is computed (two
// operand sets in turn), and the steps are separated by LDS-only barriers.
// (Loading two steps ahead measured no faster: the loads are hidden, and a
// step's time is its chain of LDS exchange, division and barrier.)
// Each sweep ends with the workgroup maximum of |corr| behind a full
// barrier, which also makes the sweep's x visible to the next.
template <int CW>
__global__ void __launch_bounds__(kMaxThreads)
KGsWavefront(int nx, int ny, Rows a, const double* __restrict__ k, double* __restrict__ x,
             double tol, int limit, double omega, unsigned long long* __restrict__ info) {
  extern __shared__ double lds[];  // [2][T] west values, then reduction scratch
  double* red = lds + 2 * blockDim.x;
  const int T = blockDim.x;
  const int t = threadIdx.x;
  const int c0 = t * CW;
  const int nc = max(0, min(nx, c0 + CW) - c0);
  for (int r = t; r < nx * ny; r += T) x[r] = 0.;
  __syncthreads();
  double south[CW];
  long iter = 0;
  double diff = 0.;
  const int nsteps = ny + T - 1;
  GsRow<CW> ra, rb;
  while (true) {
    double dmax = 0.;
    // step `step` with the operands of its row in `o`
    auto Step = [&](int step, const GsRow<CW>& o) {
      const int j = step - t;
      const int par = step & 1;
      if (nc > 0 && j >= 0 && j < ny) {
        double west = (t > 0) ? lds[(par ^ 1) * T + t - 1] : 0.;
#pragma unroll
        for (int m = 0; m < CW; ++m) {
          if (m >= nc) break;
          const int i = c0 + m, r = i + nx * j;
          double sum = 0.;
          if (j > 0) sum += o.s[m] * south[m];
          if (i > 0) sum += o.w[m] * west;
          if (i < nx - 1) sum += o.e[m] * o.xe[m];
          if (j < ny - 1) sum += o.n[m] * o.xn[m];
          const double xo = o.xo[m];
          const double value = -(o.k[m] + sum) / o.c[m];
          const double corr = value - xo;
          const double ac = fabs(corr);
          dmax = (dmax < ac) ? ac : dmax;  // std::max(diff, |corr|)
          const double xn = xo + corr * omega;
          x[r] = xn;
          south[m] = xn;
          west = xn;
        }
        lds[par * T + t] = west;
      }
      LdsBarrier();
    };
    GsLoad<CW>(ra, nx, ny, c0, -t, a, k, x);
    for (int step = 0; step < nsteps; step += 2) {
      GsLoad<CW>(rb, nx, ny, c0, step + 1 - t, a, k, x);
      Step(step, ra);
      if (step + 1 < nsteps) {
        GsLoad<CW>(ra, nx, ny, c0, step + 2 - t, a, k, x);
        Step(step + 1, rb);
      }
    }
    __syncthreads();
    // workgroup max of |corr|
    double m = WaveMax(dmax);
    if ((t & 63) == 0) red[t >> 6] = m;
    __syncthreads();
    m = red[0];
    for (int q = 1; q < T / 64; ++q) m = fmax(m, red[q]);
    __syncthreads();
    diff = m;
    // do { sweep } while (diff > tol && iter++ < limit)
    if (!(diff > tol)) break;
    if (!(iter++ < limit)) break;
  }
  if (t == 0) {
    info[0] = (unsigned long long)iter;
    info[1] = (unsigned long long)__double_as_longlong(diff);
  }
}

// LuDecompositionRelaxed::Solve (linear.hpp:592-648), every iteration in
// this workgroup.  An iteration is a forward sweep on the correction
//   corr = (-f - s corr_S - w corr_W) / (c + relax)        (raw order: wavefront)
// then a backward step that reads the PREVIOUS iterate res, not corr, as the
// reference does, so it has no recurrence:
//   corr -= (n res_N + e res_E) / (c + relax)
// then res += corr with diff = max|corr|, and f = the rows evaluated at res
// (Expression::Evaluate: constant, then S, W, C, E, N).  f starts as the
// constants and res at zero.  Stops as do { ... } while (diff > tol &&
// iter++ < limit).
template <int CW>
__global__ void __launch_bounds__(kMaxThreads)
KLuRelaxed(int nx, int ny, Rows a, const double* __restrict__ k,
           double* __restrict__ res, double* __restrict__ corr, double* __restrict__ f,
           double tol, int limit, double relax, unsigned long long* __restrict__ info) {
  extern __shared__ double lds[];  // [2][T] west values, then reduction scratch
  double* red = lds + 2 * blockDim.x;
  const int T = blockDim.x;
  const int t = threadIdx.x;
  const int c0 = t * CW;
  const int nc = max(0, min(nx, c0 + CW) - c0);
  const int N = nx * ny;
  const int nsteps = ny + T - 1;
  // the forward sweep's row operands (s, w, c and f), one step ahead as in
  // the LU sweep: f changes only between sweeps, behind full barriers
  const double* const ff[2] = {f, f};
  double* const cc[2] = {corr, corr};
  LuRow<1, CW> ra, rb;
  for (int r = t; r < N; r += T) {
    res[r] = 0.;
    f[r] = k[r];
  }
  __syncthreads();
  double south[CW];
  long iter = 0;
  double diff = 0.;
  auto Fwd = [&](int step, const LuRow<1, CW>& o) {
    const int j = step - t;
    const int par = step & 1;
    if (nc > 0 && j >= 0 && j < ny) {
      double west = (t > 0) ? lds[(par ^ 1) * T + t - 1] : 0.;
#pragma unroll
      for (int m = 0; m < CW; ++m) {
        if (m >= nc) break;
        const int i = c0 + m, r = i + nx * j;
        double sum = 0;
        if (j > 0) sum += o.a[m] * south[m];
        if (i > 0) sum += o.b[m] * west;
        const double v = (-o.v[0][m] - sum) / (o.c[m] + relax);
        corr[r] = v;
        south[m] = v;
        west = v;
      }
      lds[par * T + t] = west;
    }
    LdsBarrier();
  };
  while (true) {
    // ---- forward (wavefront over i + j)
    LuLoad<1, CW, true>(ra, nx, ny, c0, -t, a, ff, cc);
    for (int step = 0; step < nsteps; step += 2) {
      LuLoad<1, CW, true>(rb, nx, ny, c0, step + 1 - t, a, ff, cc);
      Fwd(step, ra);
      if (step + 1 < nsteps) {
        LuLoad<1, CW, true>(ra, nx, ny, c0, step + 2 - t, a, ff, cc);
        Fwd(step + 1, rb);
      }
    }
    __syncthreads();  // every thread's corrections before the backward step reads them
    // ---- backward, from the previous iterate (no recurrence)
    for (int r = t; r < N; r += T) {
      const int i = r % nx, j = r / nx;
      double sum = 0;
      if (j < ny - 1) sum += a.n[r] * res[r + nx];
      if (i < nx - 1) sum += a.e[r] * res[r + 1];
      corr[r] -= sum / (a.c[r] + relax);
    }
    __syncthreads();
    double dmax = 0.;
    for (int r = t; r < N; r += T) {
      const double cr = corr[r];
      res[r] += cr;
      const double ac = fabs(cr);
      dmax = (dmax < ac) ? ac : dmax;  // std::max(diff, |corr|)
    }
    __syncthreads();
    for (int r = t; r < N; r += T) {
      const int i = r % nx, j = r / nx;
      double v = k[r];
      if (j > 0) v += res[r - nx] * a.s[r];
      if (i > 0) v += res[r - 1] * a.w[r];
      v += res[r] * a.c[r];
      if (i < nx - 1) v += res[r + 1] * a.e[r];
      if (j < ny - 1) v += res[r + nx] * a.n[r];
      f[r] = v;
    }
    // workgroup max of |corr| (its barriers also order f before the next
    // forward sweep reads it)
    double m = WaveMax(dmax);
    if ((t & 63) == 0) red[t >> 6] = m;
    __syncthreads();
    m = red[0];
    for (int q = 1; q < T / 64; ++q) m = fmax(m, red[q]);
    __syncthreads();
    diff = m;
    if (!(diff > tol)) break;
    if (!(iter++ < limit)) break;
  }
  if (t == 0) {
    info[0] = (unsigned long long)iter;
    info[1] = (unsigned long long)__double_as_longlong(diff);
  }
}

// One Jacobi sweep (linear.hpp:757-776) with a device-side stop test run by
// the last block to finish; later launches see `done` and return.
__global__ void __launch_bounds__(256)
KJacobiSweep(int nx, int ny, Rows a, const double* __restrict__ k, JacobiWork w, int sweep,
             double tol, int limit, double omega) {
  volatile int* state = w.state;
  if (state[0]) return;
  const double* __restrict__ x = (sweep & 1) ? w.xb : w.xa;
  double* __restrict__ nxt = (sweep & 1) ? w.xa : w.xb;
  double dmax = 0.;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nx * ny; r += gridDim.x * blockDim.x) {
    const int i = r % nx, j = r / nx;
    double sum = 0.;
    if (j > 0) sum += a.s[r] * x[r - nx];
    if (i > 0) sum += a.w[r] * x[r - 1];
    if (i < nx - 1) sum += a.e[r] * x[r + 1];
    if (j < ny - 1) sum += a.n[r] * x[r + nx];
    const double value = -(k[r] + sum) / a.c[r];
    const double corr = value - x[r];
    const double ac = fabs(corr);
    dmax = (dmax < ac) ? ac : dmax;
    nxt[r] = x[r] + corr * omega;
  }
  double m = WaveMax(dmax);
  __shared__ double red[4];
  __shared__ int last;
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    atomicMax(&w.diff[sweep], (unsigned long long)__double_as_longlong(m));
    __threadfence();
    last = (atomicAdd(&w.state[1], 1) == int(gridDim.x) - 1);
  }
  __syncthreads();
  if (last && threadIdx.x == 0) {
    __threadfence();
    const double diff = __longlong_as_double(
        (long long)atomicMax(&w.diff[sweep], 0ull));
    w.state[1] = 0;
    // do { sweep } while (diff > tol && iter++ < limit), iter == sweep here
    bool cont = (diff > tol) && (sweep < limit);
    w.state[2] = (diff > tol) ? sweep + 1 : sweep;  // reference's printed iter
    w.state[3] = (sweep + 1) & 1;                   // buffer holding the result
    if (!cont) state[0] = 1;
    __threadfence();
  }
}

// "lu_rb", first kernel: black cells ((i+j) odd).  Red neighbours are the
// forward red values -k/c computed on the fly.  T: the rows' storage and
// arithmetic type (float in the mixed path: the correction is solved in FP32
// from a residual formed in FP64).
template <class T>
__global__ void __launch_bounds__(256)
KLuRbBlack(int nx, int ny, RowsT<T> a, const T* __restrict__ k0, const T* __restrict__ k1,
           T* __restrict__ xb0, T* __restrict__ xb1) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nx * ny) return;
  const int i = r % nx, j = r / nx;
  if (((i + j) & 1) == 0) return;
  // neighbours at clamped indices, all loaded before use (a missing
  // neighbour's values are not used)
  const int qs = j > 0 ? r - nx : r, qw = i > 0 ? r - 1 : r;
  const int qe = i < nx - 1 ? r + 1 : r, qn = j < ny - 1 ? r + nx : r;
  const T as = a.s[r], aw = a.w[r], ae = a.e[r], an = a.n[r], ac = a.c[r];
  const T k0r = k0[r], k1r = k1[r];
  const T k0s = k0[qs], k1s = k1[qs], cs = a.c[qs];
  const T k0w = k0[qw], k1w = k1[qw], cw = a.c[qw];
  const T k0e = k0[qe], k1e = k1[qe], ce = a.c[qe];
  const T k0n = k0[qn], k1n = k1[qn], cn = a.c[qn];
  const T z = T(0);
  T s0 = z, s1 = z;
  if (j > 0) {
    s0 += as * ((-k0s - z) / cs);
    s1 += as * ((-k1s - z) / cs);
  }
  if (i > 0) {
    s0 += aw * ((-k0w - z) / cw);
    s1 += aw * ((-k1w - z) / cw);
  }
  if (i < nx - 1) {
    s0 += ae * ((-k0e - z) / ce);
    s1 += ae * ((-k1e - z) / ce);
  }
  if (j < ny - 1) {
    s0 += an * ((-k0n - z) / cn);
    s1 += an * ((-k1n - z) / cn);
  }
  xb0[r] = (-k0r - s0) / ac;
  xb1[r] = (-k1r - s1) / ac;
}

// "lu_rb", second kernel: backward update of red cells from black
// neighbours (N, E, W, S order) and the delta-form update for all cells,
// out = base + x in FP64.
template <class T>
__global__ void __launch_bounds__(256)
KLuRbRed(int nx, int ny, RowsT<T> a, const T* __restrict__ k0, const T* __restrict__ k1,
         const T* __restrict__ xb0, const T* __restrict__ xb1,
         const double* __restrict__ base0, const double* __restrict__ base1,
         double* __restrict__ out0, double* __restrict__ out1) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nx * ny) return;
  const int i = r % nx, j = r / nx;
  if ((i + j) & 1) {
    out0[r] = base0[r] + double(xb0[r]);
    out1[r] = base1[r] + double(xb1[r]);
    return;
  }
  // neighbours at clamped indices, all loaded before use
  const int qs = j > 0 ? r - nx : r, qw = i > 0 ? r - 1 : r;
  const int qe = i < nx - 1 ? r + 1 : r, qn = j < ny - 1 ? r + nx : r;
  const T as = a.s[r], aw = a.w[r], ae = a.e[r], an = a.n[r], ac = a.c[r];
  const double b0 = base0[r], b1 = base1[r];
  const T x0n = xb0[qn], x1n = xb1[qn], x0e = xb0[qe], x1e = xb1[qe];
  const T x0w = xb0[qw], x1w = xb1[qw], x0s = xb0[qs], x1s = xb1[qs];
  const T z = T(0);
  T x0 = (-k0[r] - z) / ac;
  T x1 = (-k1[r] - z) / ac;
  T s0 = z, s1 = z;
  if (j < ny - 1) { s0 += an * x0n; s1 += an * x1n; }
  if (i < nx - 1) { s0 += ae * x0e; s1 += ae * x1e; }
  if (i > 0) { s0 += aw * x0w; s1 += aw * x1w; }
  if (j > 0) { s0 += as * x0s; s1 += as * x1s; }
  x0 -= s0 / ac;
  x1 -= s1 / ac;
  out0[r] = b0 + double(x0);
  out1[r] = b1 + double(x1);
}

}  // namespace

void LaunchLuWavefront(hipStream_t st, int nx, int ny, Rows a, const double* k0,
                       const double* k1, double* x0, double* x1, const double* base0,
                       const double* base1, double* out0, double* out1) {
  int cw;
  int T = ThreadsFor(nx, &cw);
  auto launch = [&](auto nr, auto cwc) {
    constexpr int NR = decltype(nr)::value, CW = decltype(cwc)::value;
    const size_t lds = size_t(2) * NR * T * sizeof(double);
    KLuWavefront<NR, CW><<<1, T, lds, st>>>(nx, ny, a, k0, NR == 2 ? k1 : k0, x0,
                                            NR == 2 ? x1 : x0, base0, NR == 2 ? base1 : base0,
                                            out0, NR == 2 ? out1 : out0);
  };
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I4 = std::integral_constant<int, 4>;
  if (k1) {
    switch (cw) {
      case 1: launch(I2{}, I1{}); break;
      case 2: launch(I2{}, I2{}); break;
      case 3: launch(I2{}, I3{}); break;
      default: launch(I2{}, I4{}); break;
    }
  } else {
    switch (cw) {
      case 1: launch(I1{}, I1{}); break;
      case 2: launch(I1{}, I2{}); break;
      case 3: launch(I1{}, I3{}); break;
      default: launch(I1{}, I4{}); break;
    }
  }
}

void LaunchGsWavefront(hipStream_t st, int nx, int ny, Rows a, const double* k, double* x,
                       double tol, int limit, double omega, unsigned long long* info) {
  int cw;
  int T = ThreadsFor(nx, &cw);
  size_t lds = size_t(2) * T * sizeof(double) + 16 * sizeof(double);
  switch (cw) {
    case 1: KGsWavefront<1><<<1, T, lds, st>>>(nx, ny, a, k, x, tol, limit, omega, info); break;
    case 2: KGsWavefront<2><<<1, T, lds, st>>>(nx, ny, a, k, x, tol, limit, omega, info); break;
    case 3: KGsWavefront<3><<<1, T, lds, st>>>(nx, ny, a, k, x, tol, limit, omega, info); break;
    default: KGsWavefront<4><<<1, T, lds, st>>>(nx, ny, a, k, x, tol, limit, omega, info); break;
  }
}

void LaunchLuRelaxed(hipStream_t st, int nx, int ny, Rows a, const double* k, double* res,
                     double* corr, double* f, double tol, int limit, double relax,
                     unsigned long long* info) {
  int cw;
  int T = ThreadsFor(nx, &cw);
  size_t lds = size_t(2) * T * sizeof(double) + 16 * sizeof(double);
  switch (cw) {
    case 1: KLuRelaxed<1><<<1, T, lds, st>>>(nx, ny, a, k, res, corr, f, tol, limit, relax, info); break;
    case 2: KLuRelaxed<2><<<1, T, lds, st>>>(nx, ny, a, k, res, corr, f, tol, limit, relax, info); break;
    case 3: KLuRelaxed<3><<<1, T, lds, st>>>(nx, ny, a, k, res, corr, f, tol, limit, relax, info); break;
    default: KLuRelaxed<4><<<1, T, lds, st>>>(nx, ny, a, k, res, corr, f, tol, limit, relax, info); break;
  }
}

void LaunchJacobiSweeps(hipStream_t st, int nx, int ny, Rows a, const double* k, JacobiWork w,
                        int first_sweep, int nsweeps, double tol, int limit, double omega) {
  long n = long(nx) * ny;
  int blocks = int(std::min<long>((n + 255) / 256, 2048));
  for (int s = first_sweep; s < first_sweep + nsweeps; ++s)
    KJacobiSweep<<<blocks, 256, 0, st>>>(nx, ny, a, k, w, s, tol, limit, omega);
}

template <class T>
void LaunchLuRbT(hipStream_t st, int nx, int ny, RowsT<T> a, const T* k0, const T* k1, T* xb0,
                 T* xb1, const double* base0, const double* base1, double* out0, double* out1) {
  long n = long(nx) * ny;
  dim3 grid(unsigned((n + 255) / 256));
  KLuRbBlack<T><<<grid, 256, 0, st>>>(nx, ny, a, k0, k1, xb0, xb1);
  KLuRbRed<T><<<grid, 256, 0, st>>>(nx, ny, a, k0, k1, xb0, xb1, base0, base1, out0, out1);
}

void LaunchLuRb(hipStream_t st, int nx, int ny, Rows a, const double* k0, const double* k1,
                double* xb0, double* xb1, const double* base0, const double* base1,
                double* out0, double* out1) {
  LaunchLuRbT<double>(st, nx, ny, a, k0, k1, xb0, xb1, base0, base1, out0, out1);
}

void LaunchLuRb(hipStream_t st, int nx, int ny, RowsF a, const float* k0, const float* k1,
                float* xb0, float* xb1, const double* base0, const double* base1,
                double* out0, double* out1) {
  LaunchLuRbT<float>(st, nx, ny, a, k0, k1, xb0, xb1, base0, base1, out0, out1);
}

}  // namespace hg
