// cg.hip -- Krylov solvers for the variable-coefficient pressure system.
//
// New solvers (no reference counterpart), values of linear_solver_pressure:
//   "cg"           Jacobi-preconditioned CG, FP64 throughout (the plain
//                  FP64 path); with precision "mixed", FP32 inner CG under
//                  FP64 iterative refinement.
//   "fcg_learned"  flexible CG (Notay's FCG(1)) with FP64 Krylov vectors,
//                  residuals, dot products and updates, preconditioned by an
//                  FP32 multigrid V-cycle plus, once trained, the learned
//                  network's correction.  If the preconditioner ever breaks
//                  the descent property ((r,z) <= 0 or (p,Ap) <= 0) the solve
//                  continues with Jacobi-PCG from the current iterate, to the
//                  same tolerance.
//
// Work distribution: a persistent grid of at most 2048 workgroups walks
// 64x4 tiles (tile.h), each XCD a contiguous band.  Reductions use no
// atomics: each workgroup writes its partial sums.  The scalars of an FCG
// iteration -- beta with (r, z), alpha, the projection mean, the warm-start
// multiple -- are folded once, in a fixed order, by one 1024-thread
// workgroup (KFcgBeta, KFcgAlpha, KFoldMean, KWarmScale) into CgScalars,
// which every later kernel reads; the Jacobi-PCG kernels and the stop test
// fold the few partials they need themselves, in block order, so every
// workgroup derives the same decisions.  Partial buffers alternate by
// iteration parity so a fold never races with the next producer.  The host
// queues batches of iterations and reads the scalars rarely.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "cg.h"
#include "device_common.h"
#include "learned.h"
#include "tile.h"

namespace hg {

namespace {

constexpr int kB = kTileThreads;
// persistent grid: up to 2048 workgroups (at 1024x2048, 4 tiles each); more
// workgroups overlap more memory traffic, fewer make the per-workgroup
// folds of the block partials cheaper -- 2048 measured best
constexpr int kMaxPersist = 2048;

struct CgScalars {
  double rz, pq, alpha, beta, rr, bb, tol2, zq;
  int iter, max_iter, done, fail;
  double mean;
  double warm_s;  // scale of a warm start's x
};

#define HG_CHECK(x)                                                                  \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") +            \
                                                   hipGetErrorString(e_) + " at " #x); \
  } while (0)

// Persistent tile walk: workgroup b serves tiles of its XCD's band.
struct Walk {
  TileGrid tg;
  int nx, ny;
};

template <class F>
__device__ __forceinline__ void ForTiles(const Walk& w, F&& f) {
  const int nt = w.tg.ntx * w.tg.nty;
  const int nb = gridDim.x;                    // multiple of 8
  const int xcd = blockIdx.x % kXcds;
  const int bpx = nb / kXcds;                  // workgroups per XCD
  const int lb = blockIdx.x / kXcds;
  for (int k = lb; k < w.tg.per; k += bpx) {
    const int tile = xcd * w.tg.per + k;
    if (tile >= nt) break;
    const int tx = tile % w.tg.ntx, ty = tile / w.tg.ntx;
    const int i = tx * kTX + (threadIdx.x & (kTX - 1));
    const int j = ty * kTY + (threadIdx.x >> 6);
    if (i < w.nx && j < w.ny) f(i, j);
  }
}

// Writes this workgroup's sum of v (per quantity) to part[q * nb + block].
template <int NQ>
__device__ __forceinline__ void PutPartials(const double (&v)[NQ], double* part) {
  __shared__ double red[NQ][kB / 64];
  for (int q = 0; q < NQ; ++q) {
    const double s = WaveSum(v[q]);
    if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if (threadIdx.x < NQ) {
    const int q = threadIdx.x;
    double s = 0.;
    for (int w = 0; w < kB / 64; ++w) s += red[q][w];
    part[q * gridDim.x + blockIdx.x] = s;
  }
}

// A thread's share of a partial sum: part[t] + part[t + stride] + ... (from
// 0, in that order) over indices below n.  The first K terms are loaded at
// once, at indices clamped into the array and selected afterwards, so their
// memory round trips overlap (a loop with a load per iteration waits on
// each); the rest, if any, in a loop.
template <int K>
__device__ __forceinline__ double StridedSum(const double* part, int n, int t, int stride) {
  const int last = n > 0 ? n - 1 : 0;
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = part[min(t + k * stride, last)];
  double s = 0.;
#pragma unroll
  for (int k = 0; k < K; ++k)
    if (t + k * stride < n) s += v[k];
  for (int b = t + K * stride; b < n; b += stride) s += part[b];
  return s;
}

// Sum of `n` partials, identical in every workgroup and thread.
__device__ __forceinline__ double Fold(const double* part, int n) {
  __shared__ double red[kB / 64];
  __shared__ double tot;
  double s = StridedSum<8>(part, n, threadIdx.x, kB);
  s = WaveSum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.;
    for (int w = 0; w < kB / 64; ++w) t += red[w];
    tot = t;
  }
  __syncthreads();
  const double r = tot;
  __syncthreads();
  return r;
}

// Sums of NQ partial arrays of n entries each in one pass: the same values,
// in the same order, as NQ calls of Fold, with one pair of barriers.
template <int NQ>
__device__ __forceinline__ void FoldN(const double* const (&part)[NQ], const int (&n)[NQ],
                                      double (&out)[NQ]) {
  __shared__ double red[NQ][kB / 64];
  double s[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) s[q] = StridedSum<8>(part[q], n[q], threadIdx.x, kB);
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const double v = WaveSum(s[q]);
    if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = v;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double t = 0.;
    for (int w = 0; w < kB / 64; ++w) t += red[q][w];
    out[q] = t;
  }
  __syncthreads();
}

// (A p)_c = d_c p_c - sum_nb k_f p_nb on the face form.  The five values
// are loaded unconditionally (at clamped indices on the edges, where the
// face term is then dropped by a select) so that their loads issue together
// instead of one memory round trip per branch.
template <class T, class F>
__device__ __forceinline__ T ApplyFn(int nx, int ny, const T* __restrict__ kx,
                                     const T* __restrict__ ky, int i, int j, F&& P) {
  const T kw = kx[i + (nx + 1) * j], ke = kx[i + 1 + (nx + 1) * j];
  const T ks = ky[i + nx * j], kn = ky[i + nx * (j + 1)];
  const bool w = i > 0, e = i < nx - 1, s = j > 0, n = j < ny - 1;
  const T pc = P(i, j);
  const T pw = P(w ? i - 1 : i, j), pe = P(e ? i + 1 : i, j);
  const T ps = P(i, s ? j - 1 : j), pn = P(i, n ? j + 1 : j);
  T v = (((kw + ke) + ks) + kn) * pc;
  v = w ? v - kw * pw : v;
  v = e ? v - ke * pe : v;
  v = s ? v - ks * ps : v;
  v = n ? v - kn * pn : v;
  return v;
}

// Partial-buffer layout: upd[2][2 * nb] holds (rz, rr) of the update of
// iteration k in slot k & 1; spmv[2][nb] holds (p, q) of iteration k in
// slot k & 1; dir[2 * ndir] holds (z, q_old), (r, z) of FCG (ndir = nb when
// KFcgDir writes them, the preconditioner's grid size when its last kernel
// does).
struct Parts {
  double* upd[2];
  double* spmv[2];
  double* dir;
  double* misc;
  int nb, ndir;
};

// ---- Jacobi-PCG (T = double or float vectors) ------------------------------
// Iteration k: KSpmv reads the update partials of k-1 (stop test, beta) and
// writes (p, q) partials; KUpdate forms alpha and writes (r, z), (r, r).
template <class T>
__global__ void __launch_bounds__(kB)
KPcgSpmv(Walk w, int k, const T* __restrict__ kx, const T* __restrict__ ky,
         const T* __restrict__ z, const T* __restrict__ pold, T* __restrict__ pnew,
         T* __restrict__ q, Parts pt, CgScalars* sc) {
  if (sc->done) return;
  const int nb = pt.nb;
  // stop test and beta from the previous update (or the start partials)
  double f[3];
  FoldN<3>({pt.upd[(k + 1) & 1], pt.upd[(k + 1) & 1] + nb, pt.upd[k & 1]}, {nb, nb, nb}, f);
  const double rz_new = f[0];
  const double rr = f[1];
  const bool stop = !(rr > sc->tol2 * sc->bb) || k >= sc->max_iter || !(rr == rr);
  if (stop) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->done = 1;
      sc->iter = k;
      sc->rr = rr;
    }
    return;
  }
  double beta = 0.;
  if (k > 0) beta = rz_new / f[2];
  const T b = T(beta);
  double acc = 0.;
  ForTiles(w, [&](int i, int j) {
    auto P = [&](int ii, int jj) {
      const int cc = ii + w.nx * jj;
      return z[cc] + b * pold[cc];
    };
    const T pc = P(i, j);
    const T qv = ApplyFn<T>(w.nx, w.ny, kx, ky, i, j, P);
    const int c = i + w.nx * j;
    pnew[c] = pc;
    q[c] = qv;
    acc += double(pc) * double(qv);
  });
  double v[1] = {acc};
  PutPartials<1>(v, pt.spmv[k & 1]);
}

template <class T>
__global__ void __launch_bounds__(kB)
KPcgUpdate(Walk w, int k, const T* __restrict__ p, const T* __restrict__ q,
           const T* __restrict__ dinv, T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
           Parts pt, CgScalars* sc) {
  if (sc->done) return;
  const int nb = pt.nb;
  double f[2];
  FoldN<2>({pt.upd[(k + 1) & 1], pt.spmv[k & 1]}, {nb, nb}, f);
  const double rz = f[0];
  const double pq = f[1];
  const T a = T(rz / pq);
  double srz = 0., srr = 0.;
  ForTiles(w, [&](int i, int j) {
    const int c = i + w.nx * j;
    x[c] += a * p[c];
    const T rv = r[c] - a * q[c];
    r[c] = rv;
    const T zv = rv * dinv[c];
    z[c] = zv;
    srz += double(rv) * double(zv);
    srr += double(rv) * double(rv);
  });
  double v[2] = {srz, srr};
  PutPartials<2>(v, pt.upd[k & 1]);
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->iter = k + 1;
}

// Start: z = D^-1 r, p_old = 0, partials (r, z), (r, r) in upd slot 1
// (the slot iteration 0 reads as "previous").
template <class T>
__global__ void __launch_bounds__(kB)
KPcgStart(Walk w, const T* __restrict__ r, const T* __restrict__ dinv, T* __restrict__ z,
          T* __restrict__ pold, Parts pt) {
  double srz = 0., srr = 0.;
  ForTiles(w, [&](int i, int j) {
    const int c = i + w.nx * j;
    const T zv = r[c] * dinv[c];
    z[c] = zv;
    pold[c] = T(0);
    srz += double(r[c]) * double(zv);
    srr += double(r[c]) * double(r[c]);
  });
  double v[2] = {srz, srr};
  PutPartials<2>(v, pt.upd[1]);
}

// ---- FCG(1) with an FP32 preconditioner output -----------------------------
// Check after update k-1 (one workgroup): stop test from (r, r).
__global__ void __launch_bounds__(kB)
KFcgCheck(int k, Parts pt, CgScalars* sc) {
  if (sc->done) return;
  const double rr = Fold(pt.upd[(k + 1) & 1] + pt.nb, pt.nb);
  if (threadIdx.x == 0) {
    sc->rr = rr;
    sc->iter = k;
    if (!(rr > sc->tol2 * sc->bb) || k >= sc->max_iter || !(rr == rr) || sc->fail) sc->done = 1;
  }
}

// (z, q_old), (r, z).
// Iteration 0 has no previous direction: (z, q_old) = 0 without reading q_old.
__global__ void __launch_bounds__(kB)
KFcgDir(Walk w, int k, const float* __restrict__ z, const double* __restrict__ qold,
        const double* __restrict__ r, Parts pt, CgScalars* sc) {
  if (sc->done) return;
  double zq = 0., rz = 0.;
  ForTiles(w, [&](int i, int j) {
    const int c = i + w.nx * j;
    const double zv = double(z[c]);
    if (k > 0) zq += zv * qold[c];
    rz += r[c] * zv;
  });
  double v[2] = {zq, rz};
  PutPartials<2>(v, pt.dir);
}

// p = z + beta p_old with beta = -(z, q_old) / (p_old, q_old) (sc->beta,
// from KFcgBeta); q = A p.  Iteration 0: p = z (p_old is not read).
__global__ void __launch_bounds__(kB)
KFcgSpmv(Walk w, int k, const double* __restrict__ kx, const double* __restrict__ ky,
         const float* __restrict__ z, const double* __restrict__ pold, double* __restrict__ pnew,
         double* __restrict__ q, Parts pt, CgScalars* sc) {
  // the stop flag is tested once the first tile's loads are in flight
  const int stopped = sc->done;
  const double beta = sc->beta;  // KFcgBeta
  // p is formed once per cell of the tile and its one-cell halo, in LDS, and
  // the stencil reads it there; each thread's loads (two halo-region cells,
  // its own cell's faces) are issued together
  constexpr int HX = kTX + 2, HY = kTY + 2, NH = HX * HY;
  static_assert(NH <= 2 * kB, "two staging slots per thread");
  __shared__ double sp[HY][HX];
  const int nx = w.nx, ny = w.ny;
  const int nt = w.tg.ntx * w.tg.nty;
  const int bpx = gridDim.x / kXcds, xcd = blockIdx.x % kXcds;
  double acc = 0.;
  for (int t = blockIdx.x / kXcds; t < w.tg.per; t += bpx) {
    const int tile = xcd * w.tg.per + t;
    if (tile >= nt) break;  // uniform over the block
    const int i0 = (tile % w.tg.ntx) * kTX, j0 = (tile / w.tg.ntx) * kTY;
    double zs[2], ps[2];
    bool ins[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int h = threadIdx.x + s * kB;
      const int ii = i0 - 1 + h % HX, jj = j0 - 1 + h / HX;
      ins[s] = h < NH && ii >= 0 && ii < nx && jj >= 0 && jj < ny;
      const int cc = ins[s] ? ii + nx * jj : 0;
      zs[s] = ins[s] ? double(z[cc]) : 0.;
      ps[s] = (ins[s] && k > 0) ? pold[cc] : 0.;
    }
    const int li = threadIdx.x & (kTX - 1), lj = threadIdx.x >> 6;
    const int i = i0 + li, j = j0 + lj;
    const bool in = i < nx && j < ny;
    const int fw = in ? i + (nx + 1) * j : 0, fs = in ? i + nx * j : 0;
    const double kw = in ? kx[fw] : 0., ke = in ? kx[fw + 1] : 0.;
    const double ks = in ? ky[fs] : 0., kn = in ? ky[fs + nx] : 0.;
    if (stopped) return;  // uniform; before the block's first barrier
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int h = threadIdx.x + s * kB;
      if (h < NH) sp[h / HX][h % HX] = ins[s] ? (k > 0 ? zs[s] + beta * ps[s] : zs[s]) : 0.;
    }
    __syncthreads();
    if (in) {
      const bool hw = i > 0, he = i < nx - 1, hs = j > 0, hn = j < ny - 1;
      const double pc = sp[lj + 1][li + 1];
      double qv = (((kw + ke) + ks) + kn) * pc;
      qv = hw ? qv - kw * sp[lj + 1][li] : qv;
      qv = he ? qv - ke * sp[lj + 1][li + 2] : qv;
      qv = hs ? qv - ks * sp[lj][li + 1] : qv;
      qv = hn ? qv - kn * sp[lj + 2][li + 1] : qv;
      const int c = i + nx * j;
      pnew[c] = pc;
      q[c] = qv;
      acc += pc * qv;
    }
    __syncthreads();  // sp is rewritten by the next tile
  }
  double v[1] = {acc};
  PutPartials<1>(v, pt.spmv[k & 1]);
}

// x += alpha p, r -= alpha q with alpha = (r, z) / (p, q) as KFcgAlpha folded
// it once (from the (r, z) that KFcgBeta checked); partial (r, r).
__global__ void __launch_bounds__(kB)
KFcgUpdate(Walk w, int k, const double* __restrict__ p, const double* __restrict__ q,
           double* __restrict__ x, double* __restrict__ r, Parts pt, CgScalars* sc) {
  if (sc->done) return;  // also when KFcgAlpha found (p, q) <= 0
  const double a = sc->alpha;
  double rr = 0.;
  ForTiles(w, [&](int i, int j) {
    const int c = i + w.nx * j;
    x[c] += a * p[c];
    const double rv = r[c] - a * q[c];
    r[c] = rv;
    rr += rv * rv;
  });
  double v[2] = {0., rr};
  PutPartials<2>(v, pt.upd[k & 1]);
}

// ---- helpers --------------------------------------------------------------
__global__ void __launch_bounds__(kB)
KSumPart(Walk w, const double* __restrict__ a, Parts pt) {
  double s = 0.;
  ForTiles(w, [&](int i, int j) { s += a[i + w.nx * j]; });
  double v[1] = {s};
  PutPartials<1>(v, pt.misc);
}

__global__ void __launch_bounds__(kB)
KProject(Walk w, const double* __restrict__ a, double* __restrict__ out, Parts pt) {
  const double mean = Fold(pt.misc, pt.nb) / (double(w.nx) * double(w.ny));
  ForTiles(w, [&](int i, int j) {
    const int c = i + w.nx * j;
    out[c] = a[c] - mean;
  });
}

// r = b - A x (FP64) and its partial (r, r) into misc.
__global__ void __launch_bounds__(kB)
KResidual(Walk w, const double* __restrict__ kx, const double* __restrict__ ky,
          const double* __restrict__ b, const double* __restrict__ x, double* __restrict__ r,
          Parts pt) {
  double acc = 0.;
  ForTiles(w, [&](int i, int j) {
    const int c = i + w.nx * j;
    auto P = [&](int ii, int jj) { return x[ii + w.nx * jj]; };
    const double rv = b[c] - ApplyFn<double>(w.nx, w.ny, kx, ky, i, j, P);
    r[c] = rv;
    acc += rv * rv;
  });
  double v[1] = {acc};
  PutPartials<1>(v, pt.misc);
}

// Start of a solve from x = 0: b_ = b - mean(b) (sc->mean, which KFoldMean
// folded from the misc partials) or b_ = b; r = b_; x = 0; partial (r, r) into update slot 1, the
// (r, r) iteration 0 of FCG reads.
__global__ void __launch_bounds__(kB)
KStart(Walk w, const double* __restrict__ b, int project, double* __restrict__ bp,
       double* __restrict__ r, double* __restrict__ x, Parts pt, const CgScalars* sc) {
  const double mean = project ? sc->mean : 0.;  // KFoldMean
  double acc = 0.;
  ForTiles(w, [&](int i, int j) {
    const int c = i + w.nx * j;
    const double v = project ? b[c] - mean : b[c];
    bp[c] = v;
    r[c] = v;
    x[c] = 0.;
    acc += v * v;
  });
  double v[1] = {acc};
  PutPartials<1>(v, pt.upd[1] + pt.nb);
}

// Start from a multiple s x of the x the caller left in place, s minimising
// the residual ||b_ - s A x|| (so the start is never worse than x = 0, and
// s is about 1 when x is a good guess).  Pass 1: b_ = b - mean(b) (or b),
// r = A x (held in r until pass 2), partials (b_, A x) and (A x, A x) in
// update slot 0 (free until the first update of the solve) and (b_, b_) in
// spmv slot 0 (folded into sc->bb before the first SpMV writes it).
__global__ void __launch_bounds__(kB)
KStartWarm(Walk w, const double* __restrict__ kx, const double* __restrict__ ky,
           const double* __restrict__ b, int project, double* __restrict__ bp,
           double* __restrict__ r, const double* __restrict__ x, Parts pt,
           const CgScalars* sc) {
  const double mean = project ? sc->mean : 0.;  // KFoldMean
  double ba = 0., aa = 0., bb = 0.;
  ForTiles(w, [&](int i, int j) {
    const int c = i + w.nx * j;
    const double v = project ? b[c] - mean : b[c];
    auto P = [&](int ii, int jj) { return x[ii + w.nx * jj]; };
    const double ax = ApplyFn<double>(w.nx, w.ny, kx, ky, i, j, P);
    bp[c] = v;
    r[c] = ax;
    ba += v * ax;
    aa += ax * ax;
    bb += v * v;
  });
  double v2[2] = {ba, aa}, v1[1] = {bb};
  PutPartials<2>(v2, pt.upd[0]);
  __syncthreads();  // the partial sums' LDS is reused
  PutPartials<1>(v1, pt.spmv[0]);
}

// Pass 2: s = (b_, A x) / (A x, A x) (0 when A x = 0; sc->warm_s, folded
// by KWarmScale); x <- s x,
// r <- b_ - s A x; partial (r, r) into update slot 1.
__global__ void __launch_bounds__(kB)
KStartWarmScale(Walk w, const double* __restrict__ bp, double* __restrict__ r,
                double* __restrict__ x, Parts pt, const CgScalars* sc) {
  const double s = sc->warm_s;  // KWarmScale
  double rr = 0.;
  ForTiles(w, [&](int i, int j) {
    const int c = i + w.nx * j;
    const double rv = bp[c] - s * r[c];
    x[c] = s * x[c];
    r[c] = rv;
    rr += rv * rv;
  });
  double v[1] = {rr};
  PutPartials<1>(v, pt.upd[1] + pt.nb);
}

// Sums of NQ partial arrays by one workgroup of kWide threads, the same in
// every thread.  With `stop` (a device flag, the same for the whole
// workgroup), a set flag returns false before the first barrier; the flag
// is tested after the partials are loaded, so that the loads do not wait
// for it.  The FCG scalars are folded once per iteration this way
// rather than in every workgroup of the kernel that uses them: a few
// thousand partials, which a wide workgroup folds sooner than 256 threads.
constexpr int kWide = 1024;
template <int NQ>
__device__ __forceinline__ bool WideFold(const double* const (&part)[NQ], const int (&n)[NQ],
                                         double (&out)[NQ], const int* stop = nullptr) {
  __shared__ double red[NQ][kWide / 64];
  const int stopped = stop ? *stop : 0;
  double s[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) s[q] = StridedSum<2>(part[q], n[q], threadIdx.x, kWide);
  if (stopped) return false;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    s[q] = WaveSum(s[q]);
    if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = s[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double t = 0.;
    for (int w = 0; w < kWide / 64; ++w) t += red[q][w];
    out[q] = t;
  }
  return true;
}

// FCG(1) direction scalars of iteration k from the preconditioner's dots:
// sc->rz = (r, z) and sc->beta = -(z, q_old) / (p_old, q_old) (0 at k = 0);
// (r, z) <= 0 (preconditioner not positive on r) sets fail and done.
__global__ void __launch_bounds__(kWide)
KFcgBeta(int k, Parts pt, CgScalars* sc) {
  double f[3];
  if (!WideFold<3>({pt.dir + pt.ndir, pt.dir, pt.spmv[(k + 1) & 1]}, {pt.ndir, pt.ndir, pt.nb}, f,
                   &sc->done))
    return;
  if (threadIdx.x == 0) {
    if (f[0] > 0.) {
      sc->rz = f[0];
      sc->beta = k > 0 ? -f[1] / f[2] : 0.;
    } else {
      sc->fail = 1;
      sc->done = 1;
    }
  }
}

// alpha = (r, z) / (p, q) of iteration k into sc->alpha, for an update that
// a later kernel applies; (p, q) <= 0 sets fail and done instead.
__global__ void __launch_bounds__(kWide)
KFcgAlpha(int k, Parts pt, CgScalars* sc) {
  double f[1];
  if (!WideFold<1>({pt.spmv[k & 1]}, {pt.nb}, f, &sc->done)) return;
  if (threadIdx.x == 0) {
    if (f[0] > 0.) {
      sc->alpha = sc->rz / f[0];
    } else {
      sc->fail = 1;
      sc->done = 1;
    }
  }
}

// sc->mean = (sum of the KSumPart partials) / n, the mean a singular
// system's right-hand side is projected by.
__global__ void __launch_bounds__(kWide)
KFoldMean(Parts pt, double n, CgScalars* sc) {
  double f[1];
  WideFold<1>({pt.misc}, {pt.nb}, f);
  if (threadIdx.x == 0) sc->mean = f[0] / n;
}

// sc->warm_s = (b_, A x) / (A x, A x) from KStartWarm's partials (0 when
// A x = 0).
__global__ void __launch_bounds__(kWide)
KWarmScale(Parts pt, CgScalars* sc) {
  double f[2];
  WideFold<2>({pt.upd[0], pt.upd[0] + pt.nb}, {pt.nb, pt.nb}, f);
  if (threadIdx.x == 0) sc->warm_s = f[1] > 0. ? f[0] / f[1] : 0.;
}

// Folds `nb` partials into sc->rr (and sc->bb when set_bb).
__global__ void __launch_bounds__(kB)
KTakeRR(const double* part, int nb, CgScalars* sc, int set_bb) {
  const double rr = Fold(part, nb);
  if (threadIdx.x == 0) {
    sc->rr = rr;
    if (set_bb) sc->bb = rr;
  }
}

__global__ void __launch_bounds__(kB)
KApply(Walk w, const double* __restrict__ kx, const double* __restrict__ ky,
       const double* __restrict__ x, double* __restrict__ y) {
  ForTiles(w, [&](int i, int j) {
    auto P = [&](int ii, int jj) { return x[ii + w.nx * jj]; };
    y[i + w.nx * j] = ApplyFn<double>(w.nx, w.ny, kx, ky, i, j, P);
  });
}

// The operator from input faces times `sign`: FP64 and FP32 faces, FP64 and
// FP32 inverse diagonal.  Each cell writes its west and south faces, and the
// east / north face on the last column / row.
__global__ void __launch_bounds__(kB)
KSetOperator(Walk w, const double* __restrict__ ikx, const double* __restrict__ iky, double sign,
             double* __restrict__ kx, double* __restrict__ ky, float* __restrict__ kx32,
             float* __restrict__ ky32, double* __restrict__ dinv, float* __restrict__ dinv32) {
  ForTiles(w, [&](int i, int j) {
    const int nx = w.nx;
    const int fw = i + (nx + 1) * j, fs = i + nx * j;
    const double kw = sign * ikx[fw], ke = sign * ikx[fw + 1];
    const double ks = sign * iky[fs], kn = sign * iky[fs + nx];
    kx[fw] = kw;
    kx32[fw] = float(kw);
    ky[fs] = ks;
    ky32[fs] = float(ks);
    if (i == nx - 1) {
      kx[fw + 1] = ke;
      kx32[fw + 1] = float(ke);
    }
    if (j == w.ny - 1) {
      ky[fs + nx] = kn;
      ky32[fs + nx] = float(kn);
    }
    const double d = ((kw + ke) + ks) + kn;
    const double di = d > 0. ? 1. / d : 0.;
    dinv[i + nx * j] = di;
    dinv32[i + nx * j] = float(di);
  });
}

__global__ void __launch_bounds__(kB)
KCastDown(Walk w, const double* __restrict__ a, float* __restrict__ b, float* __restrict__ zero) {
  ForTiles(w, [&](int i, int j) {
    const int c = i + w.nx * j;
    b[c] = float(a[c]);
    zero[c] = 0.f;
  });
}

__global__ void __launch_bounds__(kB)
KAddUp(Walk w, const float* __restrict__ e, double* __restrict__ x) {
  ForTiles(w, [&](int i, int j) { x[i + w.nx * j] += double(e[i + w.nx * j]); });
}

__global__ void KSetup(CgScalars* sc, double tol2, int max_iter) {
  if (threadIdx.x == 0) {
    sc->tol2 = tol2;
    sc->max_iter = max_iter;
    sc->done = 0;
    sc->fail = 0;
    sc->iter = 0;
  }
}

// Sets the FP32 inner solve's reference norm: bb = (r, r) from upd slot 1.
__global__ void __launch_bounds__(kB)
KInnerBB(Parts pt, CgScalars* sc) {
  const double rr = Fold(pt.upd[1] + pt.nb, pt.nb);
  if (threadIdx.x == 0) sc->bb = rr;
}

}  // namespace

// ---------------------------------------------------------------------------

namespace {
struct CgImpl {
  Walk w;
  int nb;
  Parts pt;
};
}  // namespace

static Walk MakeWalk(int nx, int ny, int* nb) {
  Walk w;
  w.tg = MakeTileGrid(nx, ny);
  w.nx = nx;
  w.ny = ny;
  int tiles = w.tg.ntx * w.tg.nty;
  int b = std::min(tiles, kMaxPersist);
  b = std::max(kXcds, (b + kXcds - 1) / kXcds * kXcds);
  *nb = b;
  return w;
}

CgSolver::CgSolver(int nx, int ny, hipStream_t stream) : nx_(nx), ny_(ny), stream_(stream) {
  n_ = long(nx) * ny;
  long nfx = long(nx + 1) * ny, nfy = long(nx) * (ny + 1);
  MakeWalk(nx, ny, &blocks_);
  auto A = [](void** p, size_t bytes) { HG_CHECK(hipMalloc(p, std::max<size_t>(bytes, 8))); };
  A((void**)&kx_, nfx * 8);
  A((void**)&ky_, nfy * 8);
  A((void**)&dinv_, n_ * 8);
  A((void**)&kx32_, nfx * 4);
  A((void**)&ky32_, nfy * 4);
  A((void**)&dinv32_, n_ * 4);
  for (double** p : {&r_, &z_, &pa_, &pb_, &q_, &b_}) A((void**)p, n_ * 8);
  for (float** p : {&r32_, &z32_, &pa32_, &pb32_, &q32_, &e32_, &zprev32_}) A((void**)p, n_ * 4);
  // partials: upd[2] x 2nb, spmv[2] x nb, dir 2nb, misc nb
  A((void**)&part_, size_t(blocks_) * 10 * 8);
  HG_CHECK(hipMemset(part_, 0, size_t(blocks_) * 10 * 8));
  A(&scal_, sizeof(CgScalars));
  HG_CHECK(hipMemset(scal_, 0, sizeof(CgScalars)));
  HG_CHECK(hipHostMalloc(&host_scal_, sizeof(CgScalars)));
  HG_CHECK(hipEventCreate(&ev0_));
  HG_CHECK(hipEventCreate(&ev1_));
}

CgSolver::~CgSolver() {
  for (void* p : {(void*)kx_, (void*)ky_, (void*)dinv_, (void*)kx32_, (void*)ky32_,
                  (void*)dinv32_, (void*)r_, (void*)z_, (void*)pa_, (void*)pb_, (void*)q_,
                  (void*)b_, (void*)r32_, (void*)z32_, (void*)pa32_, (void*)pb32_,
                  (void*)q32_, (void*)e32_, (void*)zprev32_, (void*)part_, scal_})
    if (p) (void)hipFree(p);
  if (host_scal_) (void)hipHostFree(host_scal_);
  if (dirx_) (void)hipFree(dirx_);
  (void)hipEventDestroy(ev0_);
  (void)hipEventDestroy(ev1_);
}

static Parts MakeParts(double* base, int nb) {
  Parts p;
  p.nb = nb;
  p.upd[0] = base;
  p.upd[1] = base + 2 * nb;
  p.spmv[0] = base + 4 * nb;
  p.spmv[1] = base + 5 * nb;
  p.dir = base + 6 * nb;
  p.misc = base + 8 * nb;
  p.ndir = nb;
  return p;
}

void CgSolver::SetOperator(const double* d_kx, const double* d_ky, double sign) {
  int nb;
  Walk w = MakeWalk(nx_, ny_, &nb);
  KSetOperator<<<nb, kB, 0, stream_>>>(w, d_kx, d_ky, sign, kx_, ky_, kx32_, ky32_, dinv_,
                                       dinv32_);
}

void CgSolver::ReadScalars() {
  HG_CHECK(hipMemcpyAsync(host_scal_, scal_, sizeof(CgScalars), hipMemcpyDeviceToHost, stream_));
  HG_CHECK(hipStreamSynchronize(stream_));
}

void CgSolver::Apply(const double* x, double* y) {
  int nb;
  Walk w = MakeWalk(nx_, ny_, &nb);
  KApply<<<nb, kB, 0, stream_>>>(w, kx_, ky_, x, y);
}

void CgSolver::QueueFinalResidual(const double* b, const double* x, double* r) {
  int nb;
  Walk w = MakeWalk(nx_, ny_, &nb);
  Parts pt = MakeParts(part_, nb);
  KResidual<<<nb, kB, 0, stream_>>>(w, kx_, ky_, b, x, r, pt);
  KTakeRR<<<1, kB, 0, stream_>>>(pt.misc, nb, (CgScalars*)scal_, 0);
  HG_CHECK(hipMemcpyAsync(host_scal_, scal_, sizeof(CgScalars), hipMemcpyDeviceToHost, stream_));
}

double CgSolver::ResidualFp64(const double* b, const double* x, double* r) {
  QueueFinalResidual(b, x, r);
  HG_CHECK(hipStreamSynchronize(stream_));
  return ((CgScalars*)host_scal_)->rr;
}

void CgSolver::FinishDeferred(CgResult& res) {
  if (res.residual_pending) {
    const CgScalars* hs = (const CgScalars*)host_scal_;
    res.rel_residual = hs->bb > 0 ? std::sqrt(hs->rr / hs->bb) : 0.;
    res.converged = res.rel_residual <= res.tol;
    res.residual_pending = false;
  }
  float ms = 0.f;
  HG_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
  res.ms = ms;
}

CgResult CgSolver::Solve(const double* b, double* x, double tol, int max_iters, int precision,
                         bool project, Precond* pc, bool defer_final, bool warm) {
  HG_CHECK(hipEventRecord(ev0_, stream_));
  CgResult res;
  if (pc)
    res = SolveFlexible(b, x, tol, max_iters, project, pc, warm);
  else if (precision == 0)
    res = SolveFp64(b, x, tol, max_iters, project, warm);
  else
    res = SolveMixed(b, x, tol, max_iters, project, nullptr);
  HG_CHECK(hipEventRecord(ev1_, stream_));
  if (defer_final) return res;
  HG_CHECK(hipEventSynchronize(ev1_));
  FinishDeferred(res);
  return res;
}

// b_ <- b (projected to zero mean when singular); sc->bb = ||b_||^2.  Cold:
// x <- 0, r_ <- b_ (the residual of x = 0).  Warm: x as the caller left it,
// r_ <- b_ - A x.  (r, r) partials in update slot 1.
void CgSolver::Prepare(const double* b, double* x, bool project, bool warm) {
  int nb;
  Walk w = MakeWalk(nx_, ny_, &nb);
  Parts pt = MakeParts(part_, nb);
  CgScalars* sc = (CgScalars*)scal_;
  if (project) {
    KSumPart<<<nb, kB, 0, stream_>>>(w, b, pt);
    KFoldMean<<<1, kWide, 0, stream_>>>(pt, double(nx_) * double(ny_), sc);
  }
  if (warm) {
    KStartWarm<<<nb, kB, 0, stream_>>>(w, kx_, ky_, b, project ? 1 : 0, b_, r_, x, pt, sc);
    KWarmScale<<<1, kWide, 0, stream_>>>(pt, sc);
    KStartWarmScale<<<nb, kB, 0, stream_>>>(w, b_, r_, x, pt, sc);
    KTakeRR<<<1, kB, 0, stream_>>>(pt.spmv[0], nb, sc, 1);
  } else {
    KStart<<<nb, kB, 0, stream_>>>(w, b, project ? 1 : 0, b_, r_, x, pt, sc);
    KTakeRR<<<1, kB, 0, stream_>>>(pt.upd[1] + nb, nb, sc, 1);
  }
}

// Jacobi-PCG on A x = b_ continuing from the current x and r_ (FP64).
int CgSolver::JacobiPcg64(double* x, int max_iters) {
  int nb;
  Walk w = MakeWalk(nx_, ny_, &nb);
  Parts pt = MakeParts(part_, nb);
  CgScalars* sc = (CgScalars*)scal_;
  CgScalars* hs = (CgScalars*)host_scal_;
  KPcgStart<double><<<nb, kB, 0, stream_>>>(w, r_, dinv_, z_, pa_, pt);
  double* pold = pa_;
  double* pnew = pb_;
  int k = 0;
  const int batch = 32;
  while (true) {
    for (int it = 0; it < batch; ++it, ++k) {
      KPcgSpmv<double><<<nb, kB, 0, stream_>>>(w, k, kx_, ky_, z_, pold, pnew, q_, pt, sc);
      KPcgUpdate<double><<<nb, kB, 0, stream_>>>(w, k, pnew, q_, dinv_, x, r_, z_, pt, sc);
      std::swap(pold, pnew);
    }
    ReadScalars();
    if (hs->done || k >= max_iters + batch) break;
  }
  return hs->iter;
}

CgResult CgSolver::SolveFp64(const double* b, double* x, double tol, int max_iters,
                             bool project, bool warm) {
  CgScalars* sc = (CgScalars*)scal_;
  KSetup<<<1, 64, 0, stream_>>>(sc, tol * tol, max_iters);
  Prepare(b, x, project, warm);
  CgResult res;
  res.iterations = JacobiPcg64(x, max_iters);
  QueueFinalResidual(b_, x, r_);
  res.tol = tol;
  res.residual_pending = true;
  res.outer = 1;
  return res;
}

CgResult CgSolver::SolveFlexible(const double* b, double* x, double tol, int max_iters,
                                 bool project, Precond* pc, bool warm) {
  int nb;
  Walk w = MakeWalk(nx_, ny_, &nb);
  Parts pt = MakeParts(part_, nb);
  CgScalars* sc = (CgScalars*)scal_;
  CgScalars* hs = (CgScalars*)host_scal_;
  KSetup<<<1, 64, 0, stream_>>>(sc, tol * tol, max_iters);
  Prepare(b, x, project, warm);
  double* pold = pa_;
  double* pnew = pb_;
  int k = 0;
  // iterations are queued in batches between host checks of the stop flag;
  // the first batch is sized from the previous solve (solves of one
  // simulation take similar counts): one more than its count, so that the
  // iteration whose stop test ends the solve is usually in it and the host
  // waits once; iterations queued past convergence return at once
  int batch = std::max(2, std::min(64, last_fcg_iters_ + 1));
  // a preconditioner that takes over the stop test and the direction dots
  // (FcgFusion) writes the dots to dirx_, sized for its grid
  // (and, when it takes the previous iteration's update, the update's
  // (r, r) partials at dirx_ + 2 nfp)
  const int nfp = pc->FcgPartials();
  if (nfp > dirx_cap_) {
    if (dirx_) HG_CHECK(hipFree(dirx_));
    dirx_ = nullptr;
    HG_CHECK(hipMalloc(&dirx_, size_t(3) * nfp * 8));
    dirx_cap_ = nfp;
  }
  // from the second iteration on, the preconditioner's first kernel runs the
  // previous iteration's update with the alpha that KFcgAlpha folded after
  // that iteration's SpMV, and no KFcgUpdate is queued
  const bool fuse = nfp > 0 && pc->FusesUpdate();
  // the folded update reads the residual on its tiles' halos while other
  // workgroups update theirs, so it writes the new residual to the other
  // buffer of a pair (z_, unused by FCG, is the second), which the up-leg
  // then reads
  double* rcur = r_;
  double* ralt = z_;
  CgResult res;
  while (true) {
    for (int it = 0; it < batch; ++it, ++k) {
      Parts pk = pt;
      FcgFusion f{pt.upd[(k + 1) & 1] + nb, nb, k, &sc->rr, &sc->iter, &sc->done,
                  &sc->tol2, &sc->bb, &sc->max_iter, &sc->fail, k > 0 ? q_ : nullptr, dirx_};
      if (fuse && k > 0) {
        f.rr_part = dirx_ + 2 * nfp;
        f.nb = nfp;
        f.up_on = 1;
        f.up_alpha = &sc->alpha;
        f.up_p = pold;
        f.up_q = q_;
        f.up_x = x;
        f.up_r = ralt;
        f.up_rr = dirx_ + 2 * nfp;
      }
      int ndir = 0;
      if (nfp > 0 && pc->ApplyFcg(rcur, z32_, stream_, &sc->done, f, &ndir)) {
        pk.dir = dirx_;
        pk.ndir = ndir;
        if (f.up_on) std::swap(rcur, ralt);
      } else {
        KFcgCheck<<<1, kB, 0, stream_>>>(k, pt, sc);
        pc->Apply(r_, z32_, stream_, &sc->done);
        KFcgDir<<<nb, kB, 0, stream_>>>(w, k, z32_, q_, r_, pt, sc);
      }
      KFcgBeta<<<1, kWide, 0, stream_>>>(k, pk, sc);
      KFcgSpmv<<<nb, kB, 0, stream_>>>(w, k, kx_, ky_, z32_, pold, pnew, q_, pk, sc);
      KFcgAlpha<<<1, kWide, 0, stream_>>>(k, pk, sc);
      if (!fuse) KFcgUpdate<<<nb, kB, 0, stream_>>>(w, k, pnew, q_, x, r_, pk, sc);
      std::swap(pold, pnew);
    }
    ReadScalars();
    if (hs->done || k >= max_iters + batch) break;
    batch = 2;
  }
  res.iterations = hs->iter;
  last_fcg_iters_ = hs->fail ? 4 : std::max(2, hs->iter);
  if (hs->fail) {
    // the preconditioner broke the descent property: finish with
    // Jacobi-PCG from the current iterate (same FP64 tolerance; sc->bb
    // still holds ||b||^2 from Prepare)
    res.used_fallback = true;
    const int left = std::max(1, max_iters - res.iterations);
    KSetup<<<1, 64, 0, stream_>>>(sc, tol * tol, left);
    KResidual<<<nb, kB, 0, stream_>>>(w, kx_, ky_, b_, x, r_, pt);
    res.iterations += JacobiPcg64(x, left);
  }
  QueueFinalResidual(b_, x, r_);
  res.tol = tol;
  res.residual_pending = true;
  res.outer = 1;
  return res;
}

int CgSolver::InnerFp32(double tol_inner, int max_iters, Precond*, bool*) {
  int nb;
  Walk w = MakeWalk(nx_, ny_, &nb);
  Parts pt = MakeParts(part_, nb);
  CgScalars* sc = (CgScalars*)scal_;
  CgScalars* hs = (CgScalars*)host_scal_;
  KSetup<<<1, 64, 0, stream_>>>(sc, tol_inner * tol_inner, max_iters);
  KPcgStart<float><<<nb, kB, 0, stream_>>>(w, r32_, dinv32_, z32_, pa32_, pt);
  KInnerBB<<<1, kB, 0, stream_>>>(pt, sc);
  float* pold = pa32_;
  float* pnew = pb32_;
  int k = 0;
  const int batch = 32;
  while (true) {
    for (int it = 0; it < batch; ++it, ++k) {
      KPcgSpmv<float><<<nb, kB, 0, stream_>>>(w, k, kx32_, ky32_, z32_, pold, pnew, q32_, pt, sc);
      KPcgUpdate<float><<<nb, kB, 0, stream_>>>(w, k, pnew, q32_, dinv32_, e32_, r32_, z32_, pt,
                                                sc);
      std::swap(pold, pnew);
    }
    ReadScalars();
    if (hs->done || k >= max_iters + batch) break;
  }
  return hs->iter;
}

// FP32 inner Jacobi-PCG under FP64 iterative refinement ("cg", mixed).
CgResult CgSolver::SolveMixed(const double* b, double* x, double tol, int max_iters,
                              bool project, Precond*) {
  int nb;
  Walk w = MakeWalk(nx_, ny_, &nb);
  Parts pt = MakeParts(part_, nb);
  CgScalars* sc = (CgScalars*)scal_;
  CgResult res;
  KSetup<<<1, 64, 0, stream_>>>(sc, 0., 1);
  Prepare(b, x, project, false);
  ReadScalars();
  double bb = ((CgScalars*)host_scal_)->bb;
  if (bb == 0.) {
    res.converged = true;
    return res;
  }
  double rr = bb;
  int total = 0;
  while (std::sqrt(rr / bb) > tol && total < max_iters && res.outer < 50) {
    if (project) {
      KSumPart<<<nb, kB, 0, stream_>>>(w, r_, pt);
      KProject<<<nb, kB, 0, stream_>>>(w, r_, r_, pt);
    }
    KCastDown<<<nb, kB, 0, stream_>>>(w, r_, r32_, e32_);
    double need = tol * std::sqrt(bb / rr);
    double itol = std::max(1e-4, need * 0.5);
    total += InnerFp32(itol, max_iters - total, nullptr, nullptr);
    KAddUp<<<nb, kB, 0, stream_>>>(w, e32_, x);
    rr = ResidualFp64(b_, x, r_);
    res.outer += 1;
    if (!(rr == rr)) break;
  }
  res.iterations = total;
  res.rel_residual = std::sqrt(rr / bb);
  res.converged = res.rel_residual <= tol;
  return res;
}

}  // namespace hg
