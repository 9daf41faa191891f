// mg3d.hip -- flexible CG with an FP32 multigrid V-cycle on the 3D step's
// 7-point pressure rows (see mg3d.h).

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

#include "mg3d.h"

#define M3_CHECK(x)                                                                  \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") +            \
                                                   hipGetErrorString(e_) + " at " #x); \
  } while (0)

namespace hg3 {

namespace {

constexpr int kB = 256;          // threads of the level and Krylov kernels
constexpr int kMaxBlocks = 1024;  // grid of the Krylov kernels (grid-stride)
constexpr int kCoarseT = 1024;    // threads of the coarsest solve
constexpr int kCoarseSlots = Mg3Fcg::kCoarsestMax / kCoarseT;
constexpr int kCoarseIters = 24;  // Jacobi-PCG iterations on the coarsest grid (a partial solve)
constexpr long kTinyCells = 64;   // no coarsening below this size

// Chebyshev damping of the three Jacobi sweeps (reciprocal roots of T_3 on
// [1/5, 2]: r_k = 11/10 + 9/10 cos((2k-1) pi / 6)), small damping first.
constexpr float kW0 = 0.532078f, kW1 = 0.909091f, kW2 = 3.119374f;

struct Scal {
  double bb, rz, pq, pq_old, alpha, beta, rr, tol2;
  int done, fail, iters, max_iters;
};
// the rows' seven entry arrays (back south west centre east north front)
struct RowsPtr {
  const double* t[7];
};

// A level as the kernels see it: cells i + nx (j + ny k); x-faces
// i + (nx+1)(j + ny k), y-faces i + nx (j + (ny+1) k), z-faces i + nx (j + ny k).
struct Lv {
  int nx, ny, nz;
  const float *kx, *ky, *kz, *dinv;
};

__device__ __forceinline__ bool Stopped(const int* done) { return done && *done; }

__device__ __forceinline__ void Ijk(long c, int nx, int ny, int& i, int& j, int& k) {
  i = int(c % nx);
  const long t = c / nx;
  j = int(t % ny);
  k = int(t / ny);
}

// The six face coefficients of cell (i, j, k) (west east south north back front).
struct Six {
  float w, e, s, n, b, f;
};
__device__ __forceinline__ Six Faces(const Lv& L, long c, int i, int j, int k) {
  const long xw = i + long(L.nx + 1) * (j + long(L.ny) * k);
  const long ys = i + long(L.nx) * (j + long(L.ny + 1) * k);
  const long nxy = long(L.nx) * L.ny;
  return Six{L.kx[xw], L.kx[xw + 1], L.ky[ys], L.ky[ys + L.nx], L.kz[c], L.kz[c + nxy]};
}

// (A x)_c of the face form from values v(q) of cells q (no read across the
// grid's boundary, where the faces are walls).
template <class V>
__device__ __forceinline__ float ApplyAt(const Lv& L, const Six& k6, long c, int i, int j, int k,
                                         V v) {
  const long nxy = long(L.nx) * L.ny;
  float a = (((((k6.w + k6.e) + k6.s) + k6.n) + k6.b) + k6.f) * v(c);
  if (i > 0) a -= k6.w * v(c - 1);
  if (i < L.nx - 1) a -= k6.e * v(c + 1);
  if (j > 0) a -= k6.s * v(c - L.nx);
  if (j < L.ny - 1) a -= k6.n * v(c + L.nx);
  if (k > 0) a -= k6.b * v(c - nxy);
  if (k < L.nz - 1) a -= k6.f * v(c + nxy);
  return a;
}

// ---- hierarchy --------------------------------------------------------------

// Finest level from the rows: k_W = -s t_W, ... with s = sign of t_C; each
// cell writes its west, south and back faces (and the east / north / front
// one on the last column / row / layer), its inverse diagonal and s.
__global__ void __launch_bounds__(kB)
KFaces(int nx, int ny, int nz, const double* __restrict__ tb, const double* __restrict__ ts,
       const double* __restrict__ tw, const double* __restrict__ tc, const double* __restrict__ te,
       const double* __restrict__ tn, const double* __restrict__ tf, float* __restrict__ kx,
       float* __restrict__ ky, float* __restrict__ kz, float* __restrict__ dinv,
       float* __restrict__ sgn) {
  const long n = long(nx) * ny * nz;
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c >= n) return;
  int i, j, k;
  Ijk(c, nx, ny, i, j, k);
  const double s = tc[c] < 0. ? -1. : 1.;
  const float fw = float(-s * tw[c]), fe = float(-s * te[c]);
  const float fs = float(-s * ts[c]), fn = float(-s * tn[c]);
  const float fb = float(-s * tb[c]), ff = float(-s * tf[c]);
  const long xw = i + long(nx + 1) * (j + long(ny) * k);
  const long ys = i + long(nx) * (j + long(ny + 1) * k);
  const long nxy = long(nx) * ny;
  kx[xw] = fw;
  ky[ys] = fs;
  kz[c] = fb;
  if (i == nx - 1) kx[xw + 1] = fe;
  if (j == ny - 1) ky[ys + nx] = fn;
  if (k == nz - 1) kz[c + nxy] = ff;
  const float d = ((((fw + fe) + fs) + fn) + fb) + ff;
  dinv[c] = d > 0.f ? 1.f / d : 0.f;
  sgn[c] = float(s);
}

// Coarse level: each coarse face is half the sum of the four fine faces it
// covers; each coarse cell writes its west / south / back faces (and the
// far ones on the last column / row / layer) and its inverse diagonal.
__global__ void __launch_bounds__(kB)
KCoarsen(int nxf, int nyf, const float* __restrict__ kxf, const float* __restrict__ kyf,
         const float* __restrict__ kzf, int nx, int ny, int nz, float* __restrict__ kx,
         float* __restrict__ ky, float* __restrict__ kz, float* __restrict__ dinv) {
  const long n = long(nx) * ny * nz;
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c >= n) return;
  int I, J, K;
  Ijk(c, nx, ny, I, J, K);
  auto fx = [&](int i, int j, int k) { return kxf[i + long(nxf + 1) * (j + long(nyf) * k)]; };
  auto fy = [&](int i, int j, int k) { return kyf[i + long(nxf) * (j + long(nyf + 1) * k)]; };
  auto fz = [&](int i, int j, int k) { return kzf[i + long(nxf) * (j + long(nyf) * k)]; };
  auto X = [&](int i) {  // x-face at fine index i under coarse (J, K)
    return 0.5f * ((fx(i, 2 * J, 2 * K) + fx(i, 2 * J + 1, 2 * K)) +
                   (fx(i, 2 * J, 2 * K + 1) + fx(i, 2 * J + 1, 2 * K + 1)));
  };
  auto Y = [&](int j) {
    return 0.5f * ((fy(2 * I, j, 2 * K) + fy(2 * I + 1, j, 2 * K)) +
                   (fy(2 * I, j, 2 * K + 1) + fy(2 * I + 1, j, 2 * K + 1)));
  };
  auto Z = [&](int k) {
    return 0.5f * ((fz(2 * I, 2 * J, k) + fz(2 * I + 1, 2 * J, k)) +
                   (fz(2 * I, 2 * J + 1, k) + fz(2 * I + 1, 2 * J + 1, k)));
  };
  const float w = X(2 * I), e = X(2 * I + 2), s = Y(2 * J), nn = Y(2 * J + 2), b = Z(2 * K),
              f = Z(2 * K + 2);
  const long xw = I + long(nx + 1) * (J + long(ny) * K);
  const long ys = I + long(nx) * (J + long(ny + 1) * K);
  kx[xw] = w;
  ky[ys] = s;
  kz[c] = b;
  if (I == nx - 1) kx[xw + 1] = e;
  if (J == ny - 1) ky[ys + nx] = nn;
  if (K == nz - 1) kz[c + long(nx) * ny] = f;
  const float d = ((((w + e) + s) + nn) + b) + f;
  dinv[c] = d > 0.f ? 1.f / d : 0.f;
}

// ---- V-cycle kernels (one cell per thread) ----------------------------------

// The first two sweeps from zero: x1 = w0 D^-1 r (formed for the neighbours
// where A x1 needs it), out = x1 + w1 D^-1 (r - A x1).  On the finest level r
// is the FP64 residual, stored here as the level's FP32 r for the later
// kernels.
template <bool FINEST>
__global__ void __launch_bounds__(kB)
KDown12(Lv L, const double* __restrict__ r64, float* __restrict__ r32, float* __restrict__ out,
        const int* done) {
  if (Stopped(done)) return;
  const long n = long(L.nx) * L.ny * L.nz;
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c >= n) return;
  int i, j, k;
  Ijk(c, L.nx, L.ny, i, j, k);
  auto R = [&](long q) { return FINEST ? float(r64[q]) : r32[q]; };
  auto X1 = [&](long q) { return kW0 * L.dinv[q] * R(q); };
  const Six k6 = Faces(L, c, i, j, k);
  const float rc = R(c);
  const float ax = ApplyAt(L, k6, c, i, j, k, X1);
  if (FINEST) r32[c] = rc;
  out[c] = X1(c) + kW1 * L.dinv[c] * (rc - ax);
}

// One damped Jacobi sweep: out = (x + w D^-1 (r - A x)) * sgn (sgn: the
// finest level's diagonal signs on the V-cycle's last sweep, else none).
__global__ void __launch_bounds__(kB)
KSweep(Lv L, const float* __restrict__ r, const float* __restrict__ x, float w,
       const float* __restrict__ sgn, float* __restrict__ out, const int* done) {
  if (Stopped(done)) return;
  const long n = long(L.nx) * L.ny * L.nz;
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c >= n) return;
  int i, j, k;
  Ijk(c, L.nx, L.ny, i, j, k);
  const Six k6 = Faces(L, c, i, j, k);
  const float ax = ApplyAt(L, k6, c, i, j, k, [&](long q) { return x[q]; });
  float v = x[c] + w * L.dinv[c] * (r[c] - ax);
  if (sgn) v *= sgn[c];
  out[c] = v;
}

// Coarse residual: the sum of the eight children's r - A x.
__global__ void __launch_bounds__(kB)
KRestrict(Lv L, const float* __restrict__ r, const float* __restrict__ x, int nxc, int nyc,
          int nzc, float* __restrict__ rc, const int* done) {
  if (Stopped(done)) return;
  const long nc = long(nxc) * nyc * nzc;
  const long C = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (C >= nc) return;
  int I, J, K;
  Ijk(C, nxc, nyc, I, J, K);
  float sum = 0.f;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int i = 2 * I + (q & 1), j = 2 * J + ((q >> 1) & 1), k = 2 * K + (q >> 2);
    const long c = i + long(L.nx) * (j + long(L.ny) * k);
    const Six k6 = Faces(L, c, i, j, k);
    sum += r[c] - ApplyAt(L, k6, c, i, j, k, [&](long p) { return x[p]; });
  }
  rc[C] = sum;
}

// x += P e: cell-centred trilinear from the parent P and its neighbours on
// the child's side in x, y, z (and the edge and corner neighbours), each
// direction weighted by w = min(1, 4 k(P|side) / max k(P faces)) of the
// coarse face coefficients: trilinear (27, 9, 3, 1) / 64 where the couplings
// are within a factor 4, collapsing onto the parent across walls (k = 0)
// and weak couplings.
// (P e) at fine cell (i, j, k): see KProlong
__device__ __forceinline__ float Prolong(const Lv& C, const float* e, int i, int j, int k) {
  const int I = i >> 1, J = j >> 1, K = k >> 1;
  const long P = I + long(C.nx) * (J + long(C.ny) * K);
  const Six k6 = Faces(C, P, I, J, K);
  const float kmax = fmaxf(fmaxf(fmaxf(k6.w, k6.e), fmaxf(k6.s, k6.n)), fmaxf(k6.b, k6.f));
  const bool px = i & 1, py = j & 1, pz = k & 1;
  float wx = 0.f, wy = 0.f, wz = 0.f;
  if (kmax > 0.f) {
    wx = fminf(1.f, 4.f * (px ? k6.e : k6.w) / kmax);
    wy = fminf(1.f, 4.f * (py ? k6.n : k6.s) / kmax);
    wz = fminf(1.f, 4.f * (pz ? k6.f : k6.b) / kmax);
  }
  // a positive weight means an interior coarse face on that side (wall
  // faces are 0), so the neighbour (and, with two or three positive weights,
  // the edge or corner cell) exists
  const long ox = px ? 1 : -1, oy = py ? long(C.nx) : -long(C.nx),
             oz = pz ? long(C.nx) * C.ny : -long(C.nx) * C.ny;
  const float ec = e[P];
  auto D = [&](bool use, long off) { return use ? e[P + off] - ec : 0.f; };
  const bool ux = wx > 0.f, uy = wy > 0.f, uz = wz > 0.f;
  const float d1 = (wx * D(ux, ox) + wy * D(uy, oy)) + wz * D(uz, oz);
  const float d2 = (wx * wy * D(ux && uy, ox + oy) + wx * wz * D(ux && uz, ox + oz)) +
                   wy * wz * D(uy && uz, oy + oz);
  const float d3 = wx * wy * wz * D(ux && uy && uz, ox + oy + oz);
  return ec + ((9.f * d1 + 3.f * d2) + d3) * (1.f / 64.f);
}
__global__ void __launch_bounds__(kB)
KProlong(Lv C, const float* __restrict__ e, int nx, int ny, int nz, float* __restrict__ x,
         const int* done) {
  if (Stopped(done)) return;
  const long n = long(nx) * ny * nz;
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c >= n) return;
  int i, j, k;
  Ijk(c, nx, ny, i, j, k);
  x[c] += Prolong(C, e, i, j, k);
}

// Coarsest level (n <= kCoarsestMax cells): at most kCoarseIters Jacobi-PCG
// iterations in one workgroup on the right-hand side projected to zero mean
// (the all-wall face-form operator is singular), result with zero mean.
__device__ __forceinline__ float BlockSumF(float v, float* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();  // red is free: its previous readers are past this point
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  float t = 0.f;
#pragma unroll
  for (int w = 0; w < kCoarseT / 64; ++w) t += red[w];
  return t;
}

// The coarsest-level solve by the whole workgroup: r and out may be global or
// LDS; sp (kCoarsestMax floats) and red are the workgroup's LDS scratch.
__device__ __forceinline__ void CoarsestPcg(const Lv& L, const float* r, float* out, float* sp,
                                            float* red) {
  const int n = L.nx * L.ny * L.nz;
  const int t = threadIdx.x;
  Six k6[kCoarseSlots];
  float di[kCoarseSlots], rr[kCoarseSlots], x[kCoarseSlots], p[kCoarseSlots], z[kCoarseSlots];
  int ci[kCoarseSlots], cj[kCoarseSlots], ck[kCoarseSlots];
  float s = 0.f;
#pragma unroll
  for (int m = 0; m < kCoarseSlots; ++m) {
    const int c = t + m * kCoarseT;
    ci[m] = cj[m] = ck[m] = 0;
    k6[m] = Six{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    di[m] = rr[m] = x[m] = p[m] = z[m] = 0.f;
    if (c < n) {
      Ijk(c, L.nx, L.ny, ci[m], cj[m], ck[m]);
      k6[m] = Faces(L, c, ci[m], cj[m], ck[m]);
      di[m] = L.dinv[c];
      rr[m] = r[c];
      s += rr[m];
    }
  }
  const float mean = BlockSumF(s, red) / float(n);
  float rz = 0.f;
#pragma unroll
  for (int m = 0; m < kCoarseSlots; ++m) {
    if (t + m * kCoarseT >= n) continue;
    rr[m] -= mean;
    z[m] = di[m] * rr[m];
    p[m] = z[m];
    rz += rr[m] * z[m];
  }
  rz = BlockSumF(rz, red);
  const float rz0 = rz;
  for (int it = 0; it < kCoarseIters && rz > 1e-10f * rz0; ++it) {
#pragma unroll
    for (int m = 0; m < kCoarseSlots; ++m)
      if (t + m * kCoarseT < n) sp[t + m * kCoarseT] = p[m];
    __syncthreads();
    float ap[kCoarseSlots], pap = 0.f;
#pragma unroll
    for (int m = 0; m < kCoarseSlots; ++m) {
      const int c = t + m * kCoarseT;
      ap[m] = 0.f;
      if (c >= n) continue;
      ap[m] = ApplyAt(L, k6[m], c, ci[m], cj[m], ck[m], [&](long q) { return sp[q]; });
      pap += p[m] * ap[m];
    }
    pap = BlockSumF(pap, red);  // also orders this iteration's sp reads before the next writes
    if (!(pap > 0.f)) break;
    const float a = rz / pap;
    float rzn = 0.f;
#pragma unroll
    for (int m = 0; m < kCoarseSlots; ++m) {
      if (t + m * kCoarseT >= n) continue;
      x[m] += a * p[m];
      rr[m] -= a * ap[m];
      z[m] = di[m] * rr[m];
      rzn += rr[m] * z[m];
    }
    rzn = BlockSumF(rzn, red);
    const float b = rzn / rz;
#pragma unroll
    for (int m = 0; m < kCoarseSlots; ++m) p[m] = z[m] + b * p[m];
    rz = rzn;
  }
  float sx = 0.f;
#pragma unroll
  for (int m = 0; m < kCoarseSlots; ++m)
    if (t + m * kCoarseT < n) sx += x[m];
  const float xm = BlockSumF(sx, red) / float(n);
#pragma unroll
  for (int m = 0; m < kCoarseSlots; ++m)
    if (t + m * kCoarseT < n) out[t + m * kCoarseT] = x[m] - xm;
}

__global__ void __launch_bounds__(kCoarseT)
KCoarsest(Lv L, const float* __restrict__ r, float* __restrict__ out, const int* done) {
  if (Stopped(done)) return;
  __shared__ float sp[Mg3Fcg::kCoarsestMax];
  __shared__ float red[kCoarseT / 64];
  CoarsestPcg(L, r, out, sp, red);
}

// The V-cycle's levels of at most kFusedMax cells, down to the coarsest, in
// one workgroup: one launch instead of eight per level plus the coarsest
// solve, the levels being too small to fill the GPU or hide a launch.  The
// same sweeps, restriction and prolongation as the grid kernels (x1 = w0
// D^-1 r stored instead of formed on the fly: the same values), on
// residuals and iterates in LDS; the faces are read from global memory,
// where these small levels stay cached.  The coarsest level gets
// CoarsestPcg.  Level 0 of the set reads its residual from fs.r and writes
// its correction to fs.out.
constexpr int kFusedMax = 4096;
constexpr int kFusedCells = 4688;  // a chain from 4096 cells by eighths: 4096 + 512 + 64 + 8 + 1
// levels halve while they hold more than 64 cells, so a set starting at <=
// 4096 cells smooths at most two levels (4096, 512) above its coarsest
constexpr int kMaxFused = 3;
struct Fused3 {
  int count;  // levels; the last is the coarsest
  int nx[kMaxFused], ny[kMaxFused], nz[kMaxFused], off[kMaxFused];
  const float *kx[kMaxFused], *ky[kMaxFused], *kz[kMaxFused], *dinv[kMaxFused];
  const float* r;
  float* out;
};

// cells per thread of fused level l: the first holds at most kFusedMax cells,
// each later one at most an eighth of that
constexpr int FusedSlots(int l) { return l == 0 ? kFusedMax / kCoarseT : 1; }
static_assert(kFusedMax / 8 <= kCoarseT, "one slot on the fused levels after the first");

__global__ void __launch_bounds__(kCoarseT)
KFused(Fused3 fs, const int* done) {
  const int stopped = done ? *done : 0;  // tested once the loads are in flight
  __shared__ float sr[kFusedCells], sxa[kFusedCells], sxb[kFusedCells];
  __shared__ float sp[Mg3Fcg::kCoarsestMax];
  __shared__ float red[kCoarseT / 64];
  const int t = threadIdx.x;
  const int last = fs.count - 1;
  auto Lev = [&](int l) {
    return Lv{fs.nx[l], fs.ny[l], fs.nz[l], fs.kx[l], fs.ky[l], fs.kz[l], fs.dinv[l]};
  };
  auto Cells = [&](int l) { return fs.nx[l] * fs.ny[l] * fs.nz[l]; };
  // the faces and inverse diagonal of this thread's cells on every smoothed
  // level, loaded once into registers (the phases then touch LDS only)
  Six k6[kMaxFused - 1][FusedSlots(0)];
  float di[kMaxFused - 1][FusedSlots(0)];
#pragma unroll
  for (int l = 0; l < kMaxFused - 1; ++l)
#pragma unroll
    for (int m = 0; m < FusedSlots(l); ++m) {
      k6[l][m] = Six{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      di[l][m] = 0.f;
      const int c = t + m * kCoarseT;
      if (l >= last || c >= Cells(l)) continue;
      int i, j, k;
      Ijk(c, fs.nx[l], fs.ny[l], i, j, k);
      k6[l][m] = Faces(Lev(l), c, i, j, k);
      di[l][m] = fs.dinv[l][c];
    }
  {
    const int n0 = Cells(0);
    for (int c = t; c < n0; c += kCoarseT) sr[c] = fs.r[c];
  }
  if (stopped) return;
  __syncthreads();
  // down: x1 -> xb, x2 -> xa, x3 -> xb, residual -> xa, restriction
#pragma unroll
  for (int l = 0; l < kMaxFused - 1; ++l) {
    if (l >= last) break;
    const Lv L = Lev(l);
    const int n = Cells(l);
    const float* r = sr + fs.off[l];
    float *xa = sxa + fs.off[l], *xb = sxb + fs.off[l];
    // o = x + w D^-1 (r - A x) (x null: the first sweep from zero)
    auto Sweep = [&](const float* x, float w, float* o) {
#pragma unroll
      for (int m = 0; m < FusedSlots(l); ++m) {
        const int c = t + m * kCoarseT;
        if (c >= n) continue;
        if (!x) {
          o[c] = w * di[l][m] * r[c];
          continue;
        }
        int i, j, k;
        Ijk(c, L.nx, L.ny, i, j, k);
        o[c] = x[c] + w * di[l][m] *
                          (r[c] - ApplyAt(L, k6[l][m], c, i, j, k, [&](long q) { return x[q]; }));
      }
      __syncthreads();
    };
    Sweep(nullptr, kW0, xb);
    Sweep(xb, kW1, xa);
    Sweep(xa, kW2, xb);
#pragma unroll
    for (int m = 0; m < FusedSlots(l); ++m) {
      const int c = t + m * kCoarseT;
      if (c >= n) continue;
      int i, j, k;
      Ijk(c, L.nx, L.ny, i, j, k);
      xa[c] = r[c] - ApplyAt(L, k6[l][m], c, i, j, k, [&](long q) { return xb[q]; });
    }
    __syncthreads();
    const int nxc = fs.nx[l + 1], nyc = fs.ny[l + 1], nc = Cells(l + 1);
    float* rc = sr + fs.off[l + 1];
    for (int C = t; C < nc; C += kCoarseT) {
      int I, J, K;
      Ijk(C, nxc, nyc, I, J, K);
      float sum = 0.f;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int i = 2 * I + (q & 1), j = 2 * J + ((q >> 1) & 1), k = 2 * K + (q >> 2);
        sum += xa[i + L.nx * (j + L.ny * k)];
      }
      rc[C] = sum;
    }
    __syncthreads();
  }
  CoarsestPcg(Lev(last), sr + fs.off[last], sxa + fs.off[last], sp, red);
  __syncthreads();
  // up: xb += P e, then xb -> xa -> xb -> xa (the level's correction in xa)
#pragma unroll
  for (int l = kMaxFused - 2; l >= 0; --l) {
    if (l >= last) continue;
    const Lv L = Lev(l), C = Lev(l + 1);
    const int n = Cells(l);
    const float* r = sr + fs.off[l];
    const float* e = sxa + fs.off[l + 1];
    float *xa = sxa + fs.off[l], *xb = sxb + fs.off[l];
#pragma unroll
    for (int m = 0; m < FusedSlots(l); ++m) {
      const int c = t + m * kCoarseT;
      if (c >= n) continue;
      int i, j, k;
      Ijk(c, L.nx, L.ny, i, j, k);
      xb[c] += Prolong(C, e, i, j, k);
    }
    __syncthreads();
    auto Sweep = [&](const float* x, float w, float* o) {
#pragma unroll
      for (int m = 0; m < FusedSlots(l); ++m) {
        const int c = t + m * kCoarseT;
        if (c >= n) continue;
        int i, j, k;
        Ijk(c, L.nx, L.ny, i, j, k);
        o[c] = x[c] + w * di[l][m] *
                          (r[c] - ApplyAt(L, k6[l][m], c, i, j, k, [&](long q) { return x[q]; }));
      }
      __syncthreads();
    };
    Sweep(xb, kW0, xa);
    Sweep(xa, kW1, xb);
    Sweep(xb, kW2, xa);
  }
  const int n0 = Cells(0);
  for (int c = t; c < n0; c += kCoarseT) fs.out[c] = sxa[c];
}

// ---- flexible CG (FP64) -----------------------------------------------------

#define M3_CELLS(c, n) \
  for (long c = blockIdx.x * long(blockDim.x) + threadIdx.x; c < (n); c += long(gridDim.x) * blockDim.x)

__device__ __forceinline__ double BlockSumD(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.;
  for (int w = 0; w < int(blockDim.x >> 6); ++w) t += red[w];
  return t;
}
// one block's partial into part[blockIdx.x]
__device__ __forceinline__ void PutPartial(double v, double* part) {
  __shared__ double red[kB / 64];
  const double t = BlockSumD(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}
// A thread's share of a partial sum: part[t] + part[t + stride] + ... (from
// 0, in that order) over indices below n; the first K terms are loaded at
// once (indices clamped into the array, selected afterwards) so that their
// memory round trips overlap, the rest in a loop.
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
// the sum of n partials, the same in every thread of the (kB-thread) block
__device__ __forceinline__ double FoldPartials(const double* part, int n) {
  __shared__ double red[kB / 64];
  // n <= kMaxBlocks = 4 kB
  const double v = StridedSum<4>(part, n, threadIdx.x, blockDim.x);
  return BlockSumD(v, red);
}

// x = 0, r = b; partial (b, b).
__global__ void __launch_bounds__(kB)
KStart(long n, const double* __restrict__ b, double* __restrict__ x, double* __restrict__ r,
       double* part) {
  double bb = 0.;
  M3_CELLS(c, n) {
    const double v = b[c];
    x[c] = 0.;
    r[c] = v;
    bb += v * v;
  }
  PutPartial(bb, part);
}
__global__ void __launch_bounds__(kB)
KSetup(Scal* sc, const double* part, int nb, double tol2, int max_iters) {
  const double bb = FoldPartials(part, nb);
  if (threadIdx.x == 0) {
    *sc = Scal{};
    sc->bb = bb;
    sc->tol2 = tol2;
    sc->max_iters = max_iters;
    sc->done = (bb > 0.) ? 0 : 1;
  }
}
// partials of (r, z) and (z, q_old)
__global__ void __launch_bounds__(kB)
KDots(long n, const double* __restrict__ r, const float* __restrict__ z,
      const double* __restrict__ qold, double* part_rz, double* part_zq, const int* done) {
  if (Stopped(done)) return;
  double rz = 0., zq = 0.;
  M3_CELLS(c, n) {
    const double zv = double(z[c]);
    rz += r[c] * zv;
    if (qold) zq += zv * qold[c];
  }
  PutPartial(rz, part_rz);
  __syncthreads();
  PutPartial(zq, part_zq);
}
// beta = -(z, q_old) / (p_old, q_old) (0 at k = 0) and (r, z)
__global__ void __launch_bounds__(kB)
KBeta(Scal* sc, const double* part_rz, const double* part_zq, int nb, int k) {
  if (sc->done) return;
  const double rz = FoldPartials(part_rz, nb);
  const double zq = FoldPartials(part_zq, nb);
  if (threadIdx.x == 0) {
    sc->rz = rz;
    sc->beta = k > 0 ? -zq / sc->pq_old : 0.;
    if (!(rz != 0.) || !__builtin_isfinite(sc->beta)) {
      sc->fail = 1;
      sc->done = 1;
    }
  }
}
// p = z + beta p_old; q = A p (the rows, with p of the neighbours formed the
// same way); partial (p, q)
__global__ void __launch_bounds__(kB)
KSpmv(int nx, int ny, int nz, RowsPtr t, const float* __restrict__ z,
      const double* __restrict__ pold, double* __restrict__ pnew, double* __restrict__ q,
      double* part, const Scal* sc, int k) {
  if (sc->done) return;
  const double beta = sc->beta;
  const long n = long(nx) * ny * nz, nxy = long(nx) * ny;
  auto Pv = [&](long c) { return k > 0 ? double(z[c]) + beta * pold[c] : double(z[c]); };
  double acc = 0.;
  M3_CELLS(c, n) {
    int i, j, kk;
    Ijk(c, nx, ny, i, j, kk);
    const double pc = Pv(c);
    double s = t.t[3][c] * pc;
    if (kk > 0) s += t.t[0][c] * Pv(c - nxy);
    if (j > 0) s += t.t[1][c] * Pv(c - nx);
    if (i > 0) s += t.t[2][c] * Pv(c - 1);
    if (i < nx - 1) s += t.t[4][c] * Pv(c + 1);
    if (j < ny - 1) s += t.t[5][c] * Pv(c + nx);
    if (kk < nz - 1) s += t.t[6][c] * Pv(c + nxy);
    pnew[c] = pc;
    q[c] = s;
    acc += pc * s;
  }
  PutPartial(acc, part);
}
// alpha = (r, z) / (p, q); (r, z) (p, q) <= 0 breaks descent: fail
__global__ void __launch_bounds__(kB)
KAlpha(Scal* sc, const double* part_pq, int nb) {
  if (sc->done) return;
  const double pq = FoldPartials(part_pq, nb);
  if (threadIdx.x == 0) {
    if (sc->rz * pq > 0.) {
      sc->alpha = sc->rz / pq;
      sc->pq_old = pq;
    } else {
      sc->fail = 1;
      sc->done = 1;
    }
  }
}
// x += alpha p, r -= alpha q; partial (r, r)
__global__ void __launch_bounds__(kB)
KUpdate(long n, const double* __restrict__ p, const double* __restrict__ q,
        double* __restrict__ x, double* __restrict__ r, double* part, const Scal* sc) {
  if (sc->done) return;
  const double a = sc->alpha;
  double rr = 0.;
  M3_CELLS(c, n) {
    x[c] += a * p[c];
    const double v = r[c] - a * q[c];
    r[c] = v;
    rr += v * v;
  }
  PutPartial(rr, part);
}
__global__ void __launch_bounds__(kB)
KCheck(Scal* sc, const double* part_rr, int nb) {
  if (sc->done) return;
  const double rr = FoldPartials(part_rr, nb);
  if (threadIdx.x == 0) {
    sc->rr = rr;
    sc->iters += 1;
    if (rr <= sc->tol2 * sc->bb || sc->iters >= sc->max_iters) sc->done = 1;
  }
}

inline unsigned Grid(long n) { return unsigned((n + kB - 1) / kB); }
inline unsigned KGrid(long n) { return unsigned(std::min<long>(Grid(n), kMaxBlocks)); }

}  // namespace

bool Mg3Fcg::Supports(int nx, int ny, int nz) {
  long n = long(nx) * ny * nz;
  while (n > kTinyCells && nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0) {
    nx /= 2;
    ny /= 2;
    nz /= 2;
    n = long(nx) * ny * nz;
  }
  return n <= kCoarsestMax;
}

void* Mg3Fcg::Alloc(size_t bytes) {
  void* p = nullptr;
  M3_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 16)));
  allocs_.push_back(p);
  M3_CHECK(hipMemset(p, 0, std::max<size_t>(bytes, 16)));
  return p;
}

Mg3Fcg::Mg3Fcg(int nx, int ny, int nz) {
  if (!Supports(nx, ny, nz))
    throw std::runtime_error("fcg_mg: the grid does not halve to a coarsest level of at most 4096 cells");
  int cx = nx, cy = ny, cz = nz;
  while (true) {
    Level L;
    L.nx = cx;
    L.ny = cy;
    L.nz = cz;
    L.n = long(cx) * cy * cz;
    L.kx = (float*)Alloc(size_t(cx + 1) * cy * cz * 4);
    L.ky = (float*)Alloc(size_t(cx) * (cy + 1) * cz * 4);
    L.kz = (float*)Alloc(size_t(cx) * cy * (cz + 1) * 4);
    L.dinv = (float*)Alloc(size_t(L.n) * 4);
    L.r = (float*)Alloc(size_t(L.n) * 4);
    L.xa = (float*)Alloc(size_t(L.n) * 4);
    L.xb = (float*)Alloc(size_t(L.n) * 4);
    lv_.push_back(L);
    if (!(L.n > kTinyCells && cx % 2 == 0 && cy % 2 == 0 && cz % 2 == 0)) break;
    cx /= 2;
    cy /= 2;
    cz /= 2;
  }
  // the small levels in one workgroup: from the first level below the finest
  // of at most kFusedMax cells, if the chain fits the kernel's LDS
  for (int l = 1; l < int(lv_.size()); ++l)
    if (lv_[l].n <= kFusedMax) {
      long cells = 0;
      for (int m = l; m < int(lv_.size()); ++m) cells += lv_[m].n;
      if (cells <= kFusedCells && int(lv_.size()) - l <= kMaxFused) first_fused_ = l;
      break;
    }
  const long n = lv_[0].n;
  sgn_ = (float*)Alloc(size_t(n) * 4);
  z32_ = (float*)Alloc(size_t(n) * 4);
  r_ = (double*)Alloc(size_t(n) * 8);
  pa_ = (double*)Alloc(size_t(n) * 8);
  pb_ = (double*)Alloc(size_t(n) * 8);
  q_ = (double*)Alloc(size_t(n) * 8);
  part_ = (double*)Alloc(size_t(4) * kMaxBlocks * 8);
  scal_ = Alloc(sizeof(Scal));
  M3_CHECK(hipHostMalloc(&host_scal_, sizeof(Scal), hipHostMallocDefault));
}

Mg3Fcg::~Mg3Fcg() {
  for (void* p : allocs_) (void)hipFree(p);
  if (host_scal_) (void)hipHostFree(host_scal_);
}

void Mg3Fcg::Build(double* const (&t)[7], hipStream_t st) {
  Level& F = lv_[0];
  KFaces<<<Grid(F.n), kB, 0, st>>>(F.nx, F.ny, F.nz, t[0], t[1], t[2], t[3], t[4], t[5], t[6],
                                   F.kx, F.ky, F.kz, F.dinv, sgn_);
  for (size_t l = 1; l < lv_.size(); ++l) {
    const Level& f = lv_[l - 1];
    Level& C = lv_[l];
    KCoarsen<<<Grid(C.n), kB, 0, st>>>(f.nx, f.ny, f.kx, f.ky, f.kz, C.nx, C.ny, C.nz, C.kx, C.ky,
                                       C.kz, C.dinv);
  }
}

// Level l's correction into `out` (for l = 0 the sign-corrected z): down-leg
// x2 (xa) -> x3 (xb) -> coarse residual; the coarser level's correction into
// its own xa; up-leg xb += P e, then xb -> xa -> xb -> out.  No kernel
// writes a buffer that it (or a neighbour's thread) reads.
void Mg3Fcg::Cycle(int l, float* out, hipStream_t st, const int* done) {
  Level& L = lv_[l];
  const Lv lv{L.nx, L.ny, L.nz, L.kx, L.ky, L.kz, L.dinv};
  if (l > 0 && l == first_fused_) {
    Fused3 fs{};
    fs.count = int(lv_.size()) - l;
    int off = 0;
    for (int k = 0; k < fs.count; ++k) {
      const Level& C = lv_[l + k];
      fs.nx[k] = C.nx;
      fs.ny[k] = C.ny;
      fs.nz[k] = C.nz;
      fs.off[k] = off;
      off += int(C.n);
      fs.kx[k] = C.kx;
      fs.ky[k] = C.ky;
      fs.kz[k] = C.kz;
      fs.dinv[k] = C.dinv;
    }
    fs.r = L.r;
    fs.out = out;
    KFused<<<1, kCoarseT, 0, st>>>(fs, done);
    return;
  }
  if (l == int(lv_.size()) - 1) {
    if (l == 0)  // a single-level hierarchy: its FP32 residual first
      KDown12<true><<<Grid(L.n), kB, 0, st>>>(lv, r_, L.r, L.xa, done);
    KCoarsest<<<1, kCoarseT, 0, st>>>(lv, L.r, l == 0 ? L.xb : out, done);
    if (l == 0) KSweep<<<Grid(L.n), kB, 0, st>>>(lv, L.r, L.xb, 0.f, sgn_, out, done);
    return;
  }
  if (l == 0)
    KDown12<true><<<Grid(L.n), kB, 0, st>>>(lv, r_, L.r, L.xa, done);
  else
    KDown12<false><<<Grid(L.n), kB, 0, st>>>(lv, nullptr, L.r, L.xa, done);
  KSweep<<<Grid(L.n), kB, 0, st>>>(lv, L.r, L.xa, kW2, nullptr, L.xb, done);
  Level& C = lv_[l + 1];
  KRestrict<<<Grid(C.n), kB, 0, st>>>(lv, L.r, L.xb, C.nx, C.ny, C.nz, C.r, done);
  Cycle(l + 1, C.xa, st, done);
  const Lv cv{C.nx, C.ny, C.nz, C.kx, C.ky, C.kz, C.dinv};
  KProlong<<<Grid(L.n), kB, 0, st>>>(cv, C.xa, L.nx, L.ny, L.nz, L.xb, done);
  KSweep<<<Grid(L.n), kB, 0, st>>>(lv, L.r, L.xb, kW0, nullptr, L.xa, done);
  KSweep<<<Grid(L.n), kB, 0, st>>>(lv, L.r, L.xa, kW1, nullptr, L.xb, done);
  KSweep<<<Grid(L.n), kB, 0, st>>>(lv, L.r, L.xb, kW2, l == 0 ? sgn_ : nullptr, out, done);
}

Fcg3Result Mg3Fcg::Solve(double* const (&t)[7], const double* b, double* x, double tol,
                         int max_iters, hipStream_t st) {
  const Level& F = lv_[0];
  const long n = F.n;
  const int nb = int(KGrid(n));
  double* part_a = part_;
  double* part_b = part_ + kMaxBlocks;
  double* part_c = part_ + 2 * kMaxBlocks;
  double* part_d = part_ + 3 * kMaxBlocks;
  Scal* sc = (Scal*)scal_;
  Scal* hs = (Scal*)host_scal_;
  const int* done = &sc->done;
  Build(t, st);
  KStart<<<nb, kB, 0, st>>>(n, b, x, r_, part_a);
  KSetup<<<1, kB, 0, st>>>(sc, part_a, nb, tol * tol, max_iters);
  RowsPtr rp;
  for (int q = 0; q < 7; ++q) rp.t[q] = t[q];
  double* pold = pa_;
  double* pnew = pb_;
  int k = 0;
  // iterations are queued in batches between host reads of the stop flag:
  // the first one past the previous solve's count, so that the solve
  // usually ends in it
  int batch = std::max(2, std::min(64, last_iters_ + 1));
  while (true) {
    for (int it = 0; it < batch; ++it, ++k) {
      Cycle(0, z32_, st, done);
      KDots<<<nb, kB, 0, st>>>(n, r_, z32_, k > 0 ? q_ : nullptr, part_a, part_b, done);
      KBeta<<<1, kB, 0, st>>>(sc, part_a, part_b, nb, k);
      KSpmv<<<nb, kB, 0, st>>>(F.nx, F.ny, F.nz, rp, z32_, pold, pnew, q_, part_c, sc, k);
      KAlpha<<<1, kB, 0, st>>>(sc, part_c, nb);
      KUpdate<<<nb, kB, 0, st>>>(n, pnew, q_, x, r_, part_d, sc);
      KCheck<<<1, kB, 0, st>>>(sc, part_d, nb);
      std::swap(pold, pnew);
    }
    M3_CHECK(hipMemcpyAsync(hs, sc, sizeof(Scal), hipMemcpyDeviceToHost, st));
    M3_CHECK(hipStreamSynchronize(st));
    if (hs->done || k >= max_iters + batch) break;
    batch = 2;
  }
  Fcg3Result res;
  res.iterations = hs->iters;
  res.failed = hs->fail != 0;
  res.converged = !res.failed && hs->rr <= hs->tol2 * hs->bb;
  last_iters_ = res.failed ? 8 : std::max(2, hs->iters);
  return res;
}

}  // namespace hg3
