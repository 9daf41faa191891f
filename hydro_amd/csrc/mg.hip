// mg.hip -- FP32 multigrid V-cycle preconditioner (see mg.h).

#include <algorithm>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "mg.h"

#define HG_CHECK(x)                                                                  \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") +            \
                                                   hipGetErrorString(e_) + " at " #x); \
  } while (0)

namespace hg {

namespace {

constexpr int kCoarsestMax = 1024;
constexpr int kTinyMax = 64;          // coarsest grid solved by one wave
constexpr int kInKernelMax = 2048;    // levels handled by the one-workgroup cycle
// Jacobi-PCG iterations on the coarsest grid: a partial solve.  The V-cycle
// is applied inside flexible CG, which tolerates the inexact coarse solve;
// on the 4x8 coarsest grid of the 1024x2048 hierarchy 8 iterations leave the
// outer FCG iteration counts unchanged and cut the coarse kernel from 31 to
// 26 us.
constexpr int kTinyIters = 8;

__device__ __forceinline__ bool Done(const int* d) { return d && *d; }

// A x at (i,j) for face-form coefficients.
template <class M, class T>
__device__ __forceinline__ M Ax(int nx, int ny, const M* __restrict__ kx,
                                    const M* __restrict__ ky, const T* __restrict__ x, int i,
                                    int j) {
  const int c = i + nx * j;
  const M kw = kx[i + (nx + 1) * j], ke = kx[i + 1 + (nx + 1) * j];
  const M ks = ky[i + nx * j], kn = ky[i + nx * (j + 1)];
  const M xc = M(x[c]);
  M v = (((kw + ke) + ks) + kn) * xc;
  if (i > 0) v -= kw * M(x[c - 1]);
  if (i < nx - 1) v -= ke * M(x[c + 1]);
  if (j > 0) v -= ks * M(x[c - nx]);
  if (j < ny - 1) v -= kn * M(x[c + nx]);
  return v;
}

// Coarse cell (I, J) of a level: its west and south face coefficients (and
// the east / north one on the last column / row), each the mean of the two
// fine faces it covers, and its inverse diagonal.
template <class M>
__device__ __forceinline__ void CoarsenCell(int nxf, const M* __restrict__ kxf,
                                            const M* __restrict__ kyf, int nxc, int nyc,
                                            M* __restrict__ kxc, M* __restrict__ kyc,
                                            M* __restrict__ dinvc, int I, int J) {
  auto fx = [&](int i, int j) { return kxf[i + (nxf + 1) * j]; };
  auto fy = [&](int i, int j) { return kyf[i + nxf * j]; };
  const M kw = 0.5f * (fx(2 * I, 2 * J) + fx(2 * I, 2 * J + 1));
  const M ke = 0.5f * (fx(2 * I + 2, 2 * J) + fx(2 * I + 2, 2 * J + 1));
  const M ks = 0.5f * (fy(2 * I, 2 * J) + fy(2 * I + 1, 2 * J));
  const M kn = 0.5f * (fy(2 * I, 2 * J + 2) + fy(2 * I + 1, 2 * J + 2));
  kxc[I + (nxc + 1) * J] = kw;
  kyc[I + nxc * J] = ks;
  if (I == nxc - 1) kxc[I + 1 + (nxc + 1) * J] = ke;
  if (J == nyc - 1) kyc[I + nxc * (J + 1)] = kn;
  const M d = ((kw + ke) + ks) + kn;
  dinvc[I + nxc * J] = d > 0.f ? 1.f / d : 0.f;
}

template <class M>
__global__ void __launch_bounds__(256)
KCoarsen(int nxf, const M* __restrict__ kxf, const M* __restrict__ kyf, int nxc,
         int nyc, M* __restrict__ kxc, M* __restrict__ kyc, M* __restrict__ dinvc) {
  const long n = long(nxc) * nyc;
  for (long c = blockIdx.x * long(blockDim.x) + threadIdx.x; c < n;
       c += long(gridDim.x) * blockDim.x)
    CoarsenCell(nxf, kxf, kyf, nxc, nyc, kxc, kyc, dinvc, int(c % nxc), int(c / nxc));
}

// Levels 0..count-1 of a chain of small grids: level k (>= 1) coarsened from
// level k-1, all in one workgroup, a barrier between levels.
constexpr int kMaxChain = 16;
template <class M>
struct ChainSet {
  int count;
  int nx[kMaxChain], ny[kMaxChain];
  M *kx[kMaxChain], *ky[kMaxChain], *dinv[kMaxChain];
};

template <class M>
__global__ void __launch_bounds__(1024)
KCoarsenChain(ChainSet<M> cs) {
  for (int k = 1; k < cs.count; ++k) {
    const int nxc = cs.nx[k], nyc = cs.ny[k];
    for (int c = threadIdx.x; c < nxc * nyc; c += blockDim.x)
      CoarsenCell(cs.nx[k - 1], cs.kx[k - 1], cs.ky[k - 1], nxc, nyc, cs.kx[k], cs.ky[k],
                  cs.dinv[k], c % nxc, c / nxc);
    __threadfence_block();
    __syncthreads();
  }
}

// First pre-smoothing sweep from zero: x = omega D^-1 r.
template <class RT, class M>
__global__ void __launch_bounds__(kTileThreads)
KSmoothFirst(int nx, int ny, TileGrid tg, const RT* __restrict__ r,
             const M* __restrict__ dinv, M omega, M* __restrict__ x,
             const int* done) {
  if (Done(done)) return;
  int i, j;
  if (!TileCell(tg, nx, ny, i, j)) return;
  const int c = i + nx * j;
  x[c] = omega * dinv[c] * M(r[c]);
}

// Damped Jacobi sweep: xn = x + omega D^-1 (r - A x).
template <class RT, class M>
__global__ void __launch_bounds__(kTileThreads)
KJacobi(int nx, int ny, TileGrid tg, const M* __restrict__ kx, const M* __restrict__ ky,
        const M* __restrict__ dinv, const RT* __restrict__ r, const M* __restrict__ x,
        M omega, M* __restrict__ xn, const int* done) {
  if (Done(done)) return;
  int i, j;
  if (!TileCell(tg, nx, ny, i, j)) return;
  const int c = i + nx * j;
  xn[c] = x[c] + omega * dinv[c] * (M(r[c]) - Ax(nx, ny, kx, ky, x, i, j));
}

// Coarse correction at a fine cell.  Injection: the parent's value.
// Bilinear (cell-centred) weighted by coupling: the parent P, its neighbours
// X (in x, on the child's side), Y (in y) and the diagonal D contribute
//   e = eP + (3 wx (eX - eP) + 3 wy (eY - eP) + wx wy (eD - eP)) / 16
// with wx = min(1, 4 k(P|X) / max k(P faces)) (coarse face coefficients):
// coefficients within a factor 4 of each other give the 9/16, 3/16, 3/16,
// 1/16 bilinear weights; a wall (k = 0) or a weakly coupled neighbour across
// a strong density jump gives injection in that direction.
template <class M>
__device__ __forceinline__ void SideWeights(M kw, M ke, M ks, M kn,
                                            M (&wt)[4]) {
  const M kmax = fmax(fmax(kw, ke), fmax(ks, kn));
  const bool any = kmax > 0.f;
  wt[0] = any ? fmin(1.f, 4.f * kw / kmax) : 0.f;
  wt[1] = any ? fmin(1.f, 4.f * ke / kmax) : 0.f;
  wt[2] = any ? fmin(1.f, 4.f * ks / kmax) : 0.f;
  wt[3] = any ? fmin(1.f, 4.f * kn / kmax) : 0.f;
}
template <class M>
__device__ __forceinline__ M Bilinear(M ec, M ex, M ey, M ed, M wx,
                                          M wy) {
  ex = wx > 0.f ? ex : ec;
  ey = wy > 0.f ? ey : ec;
  const M wd = wx * wy;
  ed = wd > 0.f ? ed : ec;
  return ec + (3.f * wx * (ex - ec) + 3.f * wy * (ey - ec) + wd * (ed - ec)) * 0.0625f;
}

// The halo-SW region of a 64 x TY tile split into per-thread slots: slots
// 0..NT-1 are tile cells (tile position threadIdx.x + 256 s), slot NT + d - 1
// a cell of the ring at distance d = 1..SW (threads beyond a ring's length
// have none).  A thread keeps its cells
// through the phases of a fused kernel, so the cells' coefficients and
// right-hand side are loaded once, together, into registers; LDS carries
// only the iterates that neighbours exchange.
template <int TY, int SW>
struct TileSlots {
  static constexpr int NT = TY * kTX / kTileThreads;
  static constexpr int N = NT + SW;
  static_assert(SW >= 1 && SW <= kMaxSweeps, "sweep count");
  static_assert(2 * (kTX + 2 * SW) + 2 * (TY + 2 * SW - 2) <= kTileThreads,
                "ring longer than the block");
  // ring at distance d: rows -d and TY-1+d (with corners), columns -d and
  // kTX-1+d (without); position relative to the tile origin
  __device__ static bool Ring(int r, int d, int& li, int& lj) {
    const int W = kTX + 2 * d, H = TY + 2 * d - 2;
    if (r < W) { li = r - d; lj = -d; return true; }
    r -= W;
    if (r < W) { li = r - d; lj = TY - 1 + d; return true; }
    r -= W;
    if (r < H) { li = -d; lj = r - d + 1; return true; }
    r -= H;
    if (r < H) { li = kTX - 1 + d; lj = r - d + 1; return true; }
    return false;
  }
  __device__ static bool Pos(int s, int& li, int& lj) {
    if (s < NT) {
      const int q = threadIdx.x + kTileThreads * s;
      li = q % kTX;
      lj = q / kTX;
      return true;
    }
    return Ring(threadIdx.x, s - NT + 1, li, lj);
  }
};

// One slot's cell: position, grid validity and (for the tile and the rings
// inside the outermost) its faces and inverse diagonal.
template <class M>
struct SlotCell {
  int li, lj, i, j;
  bool mine;  // this thread has a cell in the slot
  bool in;    // ... and it lies inside the grid
  M kw, ke, ks, kn, di;
};

// S: the tile's TileSlots<TY, SW>.
// Every load is issued, at indices clamped into the grid, and selected
// afterwards (zero outside the grid or the slot's use): the number of loads
// in flight is then the same on every path, so the tile waits on memory once
// instead of once per conditionally loaded slot.
template <class S, class RT, class M>
__device__ __forceinline__ void LoadSlots(SlotCell<M> (&sc)[S::N], RT (&rv)[S::N],
                                          int nx, int ny, int i0, int j0,
                                          const M* __restrict__ kx,
                                          const M* __restrict__ ky,
                                          const M* __restrict__ dinv,
                                          const RT* __restrict__ r, bool r_on_last_ring) {
  constexpr int SW = S::N - S::NT;
#pragma unroll
  for (int s = 0; s < S::N; ++s) {
    SlotCell<M>& c = sc[s];
    c.li = c.lj = 0;
    c.mine = S::Pos(s, c.li, c.lj);
    c.i = i0 + c.li;
    c.j = j0 + c.lj;
    c.in = c.mine && c.i >= 0 && c.i < nx && c.j >= 0 && c.j < ny;
    const int ic = min(max(c.i, 0), nx - 1), jc = min(max(c.j, 0), ny - 1);
    const int k = ic + nx * jc;
    // the tile and the inner rings carry faces (the outermost ring only
    // its inverse diagonal and, if asked, its right-hand side)
    const bool faces = s < S::NT + SW - 1;
    M kw = 0.f, ke = 0.f, ks = 0.f, kn = 0.f;
    if (faces) {  // a constant of the unrolled loop
      kw = kx[ic + (nx + 1) * jc];
      ke = kx[ic + 1 + (nx + 1) * jc];
      ks = ky[k];
      kn = ky[k + nx];
    }
    const M di = dinv[k];
    const RT rk = r[k];
    c.kw = (faces && c.in) ? kw : M(0);
    c.ke = (faces && c.in) ? ke : M(0);
    c.ks = (faces && c.in) ? ks : M(0);
    c.kn = (faces && c.in) ? kn : M(0);
    c.di = c.in ? di : M(0);
    rv[s] = (c.in && (s < S::NT + SW - 1 || r_on_last_ring)) ? rk : RT(0);
  }
}

// (A x) at slot cell c from LDS values x, the cell at x[lj + O][li + O].
template <int O, int XW, class M>
__device__ __forceinline__ M SlotApply(const SlotCell<M>& c, const M (*x)[XW], int nx,
                                           int ny) {
  const int a = c.lj + O, f = c.li + O;
  M v = (((c.kw + c.ke) + c.ks) + c.kn) * x[a][f];
  if (c.i > 0) v -= c.kw * x[a][f - 1];
  if (c.i < nx - 1) v -= c.ke * x[a][f + 1];
  if (c.j > 0) v -= c.ks * x[a - 1][f];
  if (c.j < ny - 1) v -= c.kn * x[a + 1][f];
  return v;
}

// The previous FCG iteration's update inside the finest down-leg
// (FcgFusion::up_on), alpha given: r - alpha q replaces the slots' residual
// (halo included, so the down-leg smooths the updated residual), and the
// tile cells write r, x += alpha p and contribute (r, r).
template <class S, class M>
__device__ __forceinline__ void FoldedUpdate(const FcgFusion& f, const SlotCell<M> (&sc)[S::N],
                                             double (&rv)[S::N], int nx) {
  double qv[S::N];
#pragma unroll
  for (int s = 0; s < S::N; ++s) qv[s] = sc[s].in ? f.up_q[sc[s].i + nx * sc[s].j] : 0.;
  const double a = *f.up_alpha;
  double rr = 0.;
#pragma unroll
  for (int s = 0; s < S::N; ++s) {
    if (!sc[s].in) continue;
    rv[s] = rv[s] - a * qv[s];
    if (s < S::NT) {
      const int k = sc[s].i + nx * sc[s].j;
      f.up_r[k] = rv[s];
      f.up_x[k] += a * f.up_p[k];
      rr += rv[s] * rv[s];
    }
  }
  __shared__ double red[kTileThreads / 64];
  rr = WaveSum(rr);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = rr;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.;
#pragma unroll
    for (int q = 0; q < kTileThreads / 64; ++q) t += red[q];
    f.up_rr[blockIdx.x] = t;
  }
}

// Fused down-leg of one level: SW damped Jacobi sweeps from zero
// (x1 = w1 D^-1 r, x_k+1 = x_k + w_k+1 D^-1 (r - A x_k)) and the restriction
// of r - A x_SW (sum of the four children).  A block of 256 threads works on
// a 64 x TY tile with an SW-cell halo (x_k on tile + SW - k + 1, in two LDS
// buffers used in turn): one launch and one pass over the level's arrays
// replace SW + 1, and all
// global loads are issued at the start (TileSlots), so a workgroup waits on
// memory once.  UP (FP64 residual only): the previous FCG update first
// (FoldedUpdate).
template <class M, class RT, int TY, int SW, bool UP = false>
__global__ void __launch_bounds__(kTileThreads)
KDown(int nx, int ny, TileGrid tg, const M* __restrict__ kx, const M* __restrict__ ky,
      const M* __restrict__ dinv, const RT* __restrict__ r, SweepWeights w,
      M* __restrict__ x2out, int nxc, int nyc, M* __restrict__ rc, const int* done,
      FcgFusion fcg, int stop_test) {
  static_assert(!UP || std::is_same<RT, double>::value, "the update is on the FP64 residual");
  // the stop flag is read with the data: its load is issued first and
  // tested once the tile's loads are in flight, so a working launch does
  // not wait on it before the rest
  const int stopped = *done;  // never null (MgPrecondT::never_done_)
  using S = TileSlots<TY, SW>;
  constexpr int H = SW, XW = kTX + 2 * H, YH = TY + 2 * H;
  __shared__ M sx[2][YH][XW], sres[TY][kTX];
  const int b = blockIdx.x;
  const int tile = (b % kXcds) * tg.per + b / kXcds;
  if (tile >= tg.ntx * tg.nty) {  // uniform over the block
    if (UP && threadIdx.x == 0) fcg.up_rr[b] = 0.;
    return;
  }
  const int i0 = (tile % tg.ntx) * kTX, j0 = (tile / tg.ntx) * TY;
  SlotCell<M> sc[S::N];
  RT rv[S::N];
  LoadSlots<S, RT, M>(sc, rv, nx, ny, i0, j0, kx, ky, dinv, r, true);
  if (stopped) return;
  if (stop_test && blockIdx.x == 0) FcgStopTest(fcg);
  if constexpr (UP) FoldedUpdate<S, M>(fcg, sc, rv, nx);
  // x1 on the tile and rings 1..H
#pragma unroll
  for (int s = 0; s < S::N; ++s)
    if (sc[s].mine) sx[0][sc[s].lj + H][sc[s].li + H] = sc[s].in ? w.w[0] * sc[s].di * M(rv[s]) : 0.f;
  __syncthreads();
  // x_k+1 on the tile and rings 1..H-k, from x_k (one ring wider)
#pragma unroll
  for (int k = 1; k < H; ++k) {
    const M(*src)[XW] = sx[(k - 1) & 1];
    M(*dst)[XW] = sx[k & 1];
#pragma unroll
    for (int s = 0; s < S::NT + H - k; ++s) {
      const SlotCell<M>& c = sc[s];
      if (!c.mine) continue;
      M xv = 0.f;
      if (c.in)
        xv = src[c.lj + H][c.li + H] +
             w.w[k] * c.di * (M(rv[s]) - SlotApply<H, XW>(c, src, nx, ny));
      dst[c.lj + H][c.li + H] = xv;
    }
    __syncthreads();
  }
  const M(*xs)[XW] = sx[(H - 1) & 1];
#pragma unroll
  for (int s = 0; s < S::NT; ++s) {
    const SlotCell<M>& c = sc[s];
    M res = 0.f;
    if (c.in) {
      x2out[c.i + nx * c.j] = xs[c.lj + H][c.li + H];
      res = M(rv[s]) - SlotApply<H, XW>(c, xs, nx, ny);
    }
    sres[c.lj][c.li] = res;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < (kTX / 2) * (TY / 2); q += kTileThreads) {
    const int ci = q % (kTX / 2), cj = q / (kTX / 2);
    const int I = i0 / 2 + ci, J = j0 / 2 + cj;
    if (I < nxc && J < nyc)
      rc[I + nxc * J] = ((sres[2 * cj][2 * ci] + sres[2 * cj][2 * ci + 1]) +
                         sres[2 * cj + 1][2 * ci]) + sres[2 * cj + 1][2 * ci + 1];
  }
}

// Fused up-leg of one level: y = x2 + P e (x2: the down-leg's smoothed
// iterate), then SW damped Jacobi sweeps y_k+1 = y_k + w_k+1 D^-1 (r - A y_k),
// the last into out, on the block's 64 x TY tile with an SW-cell halo (y_k
// on tile + SW - k, in two LDS buffers used in turn).  The coarse
// corrections and the coarse
// cells' side weights of the covering coarse region are staged in LDS; x2,
// r and the fine coefficients of the thread's cells (TileSlots) are loaded
// at the start with them.
template <class M, class OT, class RT, bool BILINEAR, int TY, int SW>
__global__ void __launch_bounds__(kTileThreads)
KUp(int nx, int ny, TileGrid tg, const M* __restrict__ kx, const M* __restrict__ ky,
    const M* __restrict__ dinv, const RT* __restrict__ r, const M* __restrict__ x2,
    const M* __restrict__ e, const M* __restrict__ kxc, const M* __restrict__ kyc,
    int nxc, int nyc, SweepWeights w, OT* __restrict__ out, const int* done, FcgFusion fcg,
    int dots) {
  const int stopped = *done;  // tested once the loads are in flight; never null
  using S = TileSlots<TY, SW>;
  constexpr int H = SW, XW = kTX + 2 * H, YH = TY + 2 * H;
  // coarse cells i0/2 - 2 .. i0/2 + kTX/2 + 1, likewise in y: the parents of
  // the tile and its halo and their neighbours on the child's side.  A child
  // i reads parents i/2 and i/2 + 1 (i odd) or i/2 - 1 (i even), so halos of
  // 2 and 3 cells (i0 - 3 odd, i0 - 2 even; i0 + kTX + 1 odd, i0 + kTX + 2
  // even) reach the same parents
  static_assert(SW >= 1 && SW <= 3, "coarse region covers halos of 1 to 3 cells");
  constexpr int CX = kTX / 2 + 4, CY = TY / 2 + 4, NC = CX * CY;
  constexpr int KC = (NC + kTileThreads - 1) / kTileThreads;
  __shared__ M sx[2][YH][XW];
  __shared__ M se[CY][CX], sw[BILINEAR ? 4 : 1][CY][CX];
  const int b = blockIdx.x;
  const int tile = (b % kXcds) * tg.per + b / kXcds;
  if (tile >= tg.ntx * tg.nty) {  // uniform over the block
    if (dots && threadIdx.x == 0) fcg.dir[b] = fcg.dir[gridDim.x + b] = 0.;
    return;
  }
  const int i0 = (tile % tg.ntx) * kTX, j0 = (tile / tg.ntx) * TY;
  const int ci0 = i0 / 2 - 2, cj0 = j0 / 2 - 2;
  // coarse loads, clamped into the coarse grid: cells outside are read only
  // behind a zero weight (walls), where the value is not used
  M ce[KC], cf[KC][4];
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    const int q = threadIdx.x + k * kTileThreads;
    const int I = min(max(ci0 + q % CX, 0), nxc - 1), J = min(max(cj0 + q / CX, 0), nyc - 1);
    const bool ok = q < NC;
    ce[k] = ok ? e[I + nxc * J] : 0.f;
    if (BILINEAR) {
      cf[k][0] = ok ? kxc[I + (nxc + 1) * J] : 0.f;
      cf[k][1] = ok ? kxc[I + 1 + (nxc + 1) * J] : 0.f;
      cf[k][2] = ok ? kyc[I + nxc * J] : 0.f;
      cf[k][3] = ok ? kyc[I + nxc * (J + 1)] : 0.f;
    }
  }
  SlotCell<M> sc[S::N];
  RT rv[S::N];
  M xv2[S::N];
  double qo[S::NT];
  LoadSlots<S, RT, M>(sc, rv, nx, ny, i0, j0, kx, ky, dinv, r, false);
#pragma unroll
  for (int s = 0; s < S::N; ++s) xv2[s] = sc[s].in ? x2[sc[s].i + nx * sc[s].j] : 0.f;
#pragma unroll
  for (int s = 0; s < S::NT; ++s)
    qo[s] = (dots && fcg.qold && sc[s].in) ? fcg.qold[sc[s].i + nx * sc[s].j] : 0.;
  if (stopped) return;
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    const int q = threadIdx.x + k * kTileThreads;
    if (q >= NC) continue;
    se[q / CX][q % CX] = ce[k];
    if (BILINEAR) {
      M wt[4];
      SideWeights(cf[k][0], cf[k][1], cf[k][2], cf[k][3], wt);
#pragma unroll
      for (int m = 0; m < 4; ++m) sw[m][q / CX][q % CX] = wt[m];
    }
  }
  __syncthreads();
  // y = x2 + P e on the tile and rings 1..H (zero outside the grid)
#pragma unroll
  for (int s = 0; s < S::N; ++s) {
    const SlotCell<M>& c = sc[s];
    if (!c.mine) continue;
    M yv = 0.f;
    if (c.in) {
      const int pi = (c.i >> 1) - ci0, pj = (c.j >> 1) - cj0;  // parent, local
      const M ec = se[pj][pi];
      M corr = ec;
      if (BILINEAR) {
        const int sx = c.i & 1, sy_ = c.j & 1;
        const M wx = sw[sx ? 1 : 0][pj][pi], wy = sw[sy_ ? 3 : 2][pj][pi];
        const int xi = pi + (sx ? 1 : -1), yj = pj + (sy_ ? 1 : -1);
        corr = Bilinear(ec, se[pj][xi], se[yj][pi], se[yj][xi], wx, wy);
      }
      yv = xv2[s] + corr;
    }
    sx[0][c.lj + H][c.li + H] = yv;
  }
  __syncthreads();
  // y_k+1 on the tile and rings 1..H-1-k, from y_k (one ring wider)
#pragma unroll
  for (int k = 0; k < H - 1; ++k) {
    const M(*src)[XW] = sx[k & 1];
    M(*dst)[XW] = sx[(k + 1) & 1];
#pragma unroll
    for (int s = 0; s < S::NT + H - 1 - k; ++s) {
      const SlotCell<M>& c = sc[s];
      if (!c.mine) continue;
      M xv = 0.f;
      if (c.in)
        xv = src[c.lj + H][c.li + H] +
             w.w[k] * c.di * (M(rv[s]) - SlotApply<H, XW>(c, src, nx, ny));
      dst[c.lj + H][c.li + H] = xv;
    }
    __syncthreads();
  }
  const M(*ys)[XW] = sx[(H - 1) & 1];
  double zq = 0., rz = 0.;
#pragma unroll
  for (int s = 0; s < S::NT; ++s) {
    const SlotCell<M>& c = sc[s];
    if (!c.in) continue;
    const M z = ys[c.lj + H][c.li + H] +
                    w.w[H - 1] * c.di * (M(rv[s]) - SlotApply<H, XW>(c, ys, nx, ny));
    const OT zo = OT(z);  // the dots use z as written
    out[c.i + nx * c.j] = zo;
    const double zd = double(zo);
    zq += zd * qo[s];
    rz += double(rv[s]) * zd;
  }
  if (dots) {  // block partials of (z, q_old) and (r, z)
    __shared__ double red[2][kTileThreads / 64];
    zq = WaveSum(zq);
    rz = WaveSum(rz);
    if ((threadIdx.x & 63) == 0) {
      red[0][threadIdx.x >> 6] = zq;
      red[1][threadIdx.x >> 6] = rz;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double a = 0., c = 0.;
      for (int q = 0; q < kTileThreads / 64; ++q) {
        a += red[0][q];
        c += red[1][q];
      }
      fcg.dir[b] = a;
      fcg.dir[gridDim.x + b] = c;
    }
  }
}

// Coarsest level: Jacobi sweeps in one workgroup, x and r in LDS.
template <class M>
__global__ void __launch_bounds__(kCoarsestMax)
KCoarsest(int nx, int ny, const M* __restrict__ kx, const M* __restrict__ ky,
          const M* __restrict__ dinv, const M* __restrict__ r, M omega, int sweeps,
          M* __restrict__ out, const int* done) {
  if (Done(done)) return;
  __shared__ M xs[2][kCoarsestMax];
  const int n = nx * ny;
  const int c = threadIdx.x;
  const bool act = c < n;
  M kw = 0, ke = 0, ks = 0, kn = 0, di = 0, rc = 0;
  int i = 0, j = 0;
  if (act) {
    i = c % nx;
    j = c / nx;
    kw = kx[i + (nx + 1) * j];
    ke = kx[i + 1 + (nx + 1) * j];
    ks = ky[i + nx * j];
    kn = ky[i + nx * (j + 1)];
    di = dinv[c];
    rc = r[c];
    xs[0][c] = omega * di * rc;
  }
  __syncthreads();
  int cur = 0;
  for (int s = 1; s < sweeps; ++s) {
    if (act) {
      const M* x = xs[cur];
      M v = (((kw + ke) + ks) + kn) * x[c];
      if (i > 0) v -= kw * x[c - 1];
      if (i < nx - 1) v -= ke * x[c + 1];
      if (j > 0) v -= ks * x[c - nx];
      if (j < ny - 1) v -= kn * x[c + nx];
      xs[cur ^ 1][c] = x[c] + omega * di * (rc - v);
    }
    __syncthreads();
    cur ^= 1;
  }
  if (act) out[c] = xs[cur][c];
}

template <class M>
__device__ __forceinline__ M WaveSumF(M v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ void BlockSync() {
  __threadfence_block();
  __syncthreads();
}

// The coarse end of the V-cycle in one workgroup of kCcThreads threads:
// levels first..last of the hierarchy (the first <= kInKernelMax cells, each
// next one a quarter of it), the same smoothing / restriction / bilinear
// prolongation as the grid kernels, and on the coarsest level (<= 64 cells,
// one cell per lane of wave 0) at most kTinyIters Jacobi-PCG iterations.
// Thread t owns cells t, t + kCcThreads, ... of every level; their face
// coefficients, inverse diagonals and prolongation weights are loaded into
// registers once, so the phases of the cycle touch only registers and LDS
// (iterates, residuals) between barriers.  Input: r of the first level
// (global); output: its correction (global, cs.out).
constexpr int kCcThreads = 1024;
constexpr int CcCells(int l) { return kInKernelMax >> (2 * l); }
constexpr int CcSlots(int l) { return CcCells(l) > kCcThreads ? CcCells(l) / kCcThreads : 1; }
constexpr int CcLevels() {
  int l = 0;
  while (CcCells(l) > kTinyMax) ++l;
  return l;
}
constexpr int kCcLevels = CcLevels();  // smoothed (non-coarsest) levels at most
constexpr int kCcSlots = CcSlots(0);
static_assert(kInKernelMax % kCcThreads == 0 || kInKernelMax < kCcThreads, "slots");

template <class M>
__global__ void __launch_bounds__(kCcThreads)
KCoarseCycle(CoarseSetT<M> cs, SweepWeights w, const int* done) {
  const int stopped = *done;  // tested once the loads are in flight; never null
  __shared__ M sxa[kLdsCells], sxb[kLdsCells], sr[kLdsCells];
  const int t = threadIdx.x;
  const int nlev = cs.count - 1;
  // per owned cell: faces w e s n, inverse diagonal, prolongation weights
  M kw[kCcLevels][kCcSlots], ke[kCcLevels][kCcSlots], ks[kCcLevels][kCcSlots],
      kn[kCcLevels][kCcSlots], di[kCcLevels][kCcSlots], wx[kCcLevels][kCcSlots],
      wy[kCcLevels][kCcSlots];
  // and its position, i in the low and j in the high 16 bits (the phases
  // would otherwise divide by the level's width in every sweep)
  int ij[kCcLevels][kCcSlots];
#pragma unroll
  for (int l = 0; l < kCcLevels; ++l) {
#pragma unroll
    for (int m = 0; m < CcSlots(l); ++m) {
      kw[l][m] = ke[l][m] = ks[l][m] = kn[l][m] = di[l][m] = wx[l][m] = wy[l][m] = 0.f;
      ij[l][m] = 0;
      if (l >= nlev) continue;
      const int nx = cs.nx[l], c = t + m * kCcThreads;
      if (c >= nx * cs.ny[l]) continue;
      const int i = c % nx, j = c / nx;
      ij[l][m] = i | (j << 16);
      kw[l][m] = cs.kx[l][i + (nx + 1) * j];
      ke[l][m] = cs.kx[l][i + 1 + (nx + 1) * j];
      ks[l][m] = cs.ky[l][i + nx * j];
      kn[l][m] = cs.ky[l][i + nx * (j + 1)];
      di[l][m] = cs.dinv[l][c];
      if (cs.bilinear) {
        // weights of the coarse neighbours on this child's side (zero across
        // walls, where the coarse face coefficient is zero)
        const int nxc = cs.nx[l + 1], I = i >> 1, J = j >> 1;
        const M cw = cs.kx[l + 1][I + (nxc + 1) * J], ce = cs.kx[l + 1][I + 1 + (nxc + 1) * J];
        const M cs_ = cs.ky[l + 1][I + nxc * J], cn = cs.ky[l + 1][I + nxc * (J + 1)];
        const M kmax = fmax(fmax(cw, ce), fmax(cs_, cn));
        if (kmax > 0.f) {
          wx[l][m] = fmin(1.f, 4.f * ((i & 1) ? ce : cw) / kmax);
          wy[l][m] = fmin(1.f, 4.f * ((j & 1) ? cn : cs_) / kmax);
        }
      }
    }
  }
  // the coarsest level's operator, one cell per lane of wave 0
  const int nxT = cs.nx[nlev], nyT = cs.ny[nlev], nT = nxT * nyT;
  const bool actT = t < nT;
  const int iT = actT ? t % nxT : 0, jT = actT ? t / nxT : 0;
  M tw = 0.f, te = 0.f, ts = 0.f, tn = 0.f, tdi = 0.f;
  if (actT) {
    tw = cs.kx[nlev][iT + (nxT + 1) * jT];
    te = cs.kx[nlev][iT + 1 + (nxT + 1) * jT];
    ts = cs.ky[nlev][iT + nxT * jT];
    tn = cs.ky[nlev][iT + nxT * (jT + 1)];
    tdi = cs.dinv[nlev][t];
  }
  if (stopped) return;
  const int n0 = cs.nx[0] * cs.ny[0];
#pragma unroll
  for (int m = 0; m < kCcSlots; ++m) {
    const int c = t + m * kCcThreads;
    if (c < n0) sr[c] = cs.r[0][c];
  }
  __syncthreads();
  // (A x) of owned cell (l, m) at (i, j) from LDS values x (level-local)
  auto Apply = [&](const M* x, int l, int m, int nx, int ny, int c, int i, int j) {
    M v = (((kw[l][m] + ke[l][m]) + ks[l][m]) + kn[l][m]) * x[c];
    if (i > 0) v -= kw[l][m] * x[c - 1];
    if (i < nx - 1) v -= ke[l][m] * x[c + 1];
    if (j > 0) v -= ks[l][m] * x[c - nx];
    if (j < ny - 1) v -= kn[l][m] * x[c + nx];
    return v;
  };
  // Sweep k (0-based) of a leg writes Buf(k): the two vectors in turn, so
  // that the last (k = kSweeps - 1) writes xb on the way down (the up-leg's
  // start) and xa on the way up (the level's result, read by the finer
  // level's prolongation).
  auto Buf = [](M* xa, M* xb, int k, bool down) {
    const bool last_parity = ((kSweeps - 1 - k) & 1) == 0;
    return (last_parity == down) ? xb : xa;
  };
  // ---- down: kSweeps sweeps from zero, residual, restriction (sum of children)
#pragma unroll
  for (int l = 0; l < kCcLevels; ++l) {
    if (l >= nlev) break;
    const int nx = cs.nx[l], ny = cs.ny[l], n = nx * ny;
    const M* r = sr + cs.off[l];
    M *xa = sxa + cs.off[l], *xb = sxb + cs.off[l];
#pragma unroll
    for (int k = 0; k < kSweeps; ++k) {
      M* dst = Buf(xa, xb, k, true);
      const M* src = Buf(xa, xb, k - 1, true);
#pragma unroll
      for (int m = 0; m < CcSlots(l); ++m) {
        const int c = t + m * kCcThreads;
        if (c >= n) continue;
        dst[c] = k == 0 ? w.w[0] * di[l][m] * r[c]
                        : src[c] + w.w[k] * di[l][m] *
                                       (r[c] - Apply(src, l, m, nx, ny, c, ij[l][m] & 0xffff,
                                                     ij[l][m] >> 16));
      }
      __syncthreads();
    }
    // residual r - A xb into xa (free until the up-leg)
#pragma unroll
    for (int m = 0; m < CcSlots(l); ++m) {
      const int c = t + m * kCcThreads;
      if (c < n) xa[c] = r[c] - Apply(xb, l, m, nx, ny, c, ij[l][m] & 0xffff, ij[l][m] >> 16);
    }
    __syncthreads();
    const int nxc = cs.nx[l + 1], nc = nxc * cs.ny[l + 1];
    M* rc = sr + cs.off[l + 1];
    for (int C = t; C < nc; C += kCcThreads) {
      const int I = C % nxc, J = C / nxc;
      M sum = 0.f;
#pragma unroll
      for (int q = 0; q < 4; ++q) sum += xa[2 * I + (q & 1) + nx * (2 * J + (q >> 1))];
      rc[C] = sum;
    }
    __syncthreads();
  }
  // ---- coarsest: Jacobi-PCG in wave 0, result into sxa
  if (t < 64) {
    const M* r = sr + cs.off[nlev];
    M b = actT ? r[t] : 0.f;
    const M dg = ((tw + te) + ts) + tn;
    // the face-form operator has the constants as null space: solve with the
    // right-hand side projected to zero mean (consistent system)
    b -= WaveSumF(b) / M(nT);
    if (!actT) b = 0.f;
    M x = 0.f, rr = b, z = tdi * rr, p = z;
    M rz = WaveSumF(rr * z);
    const M rz0 = rz;
    for (int it = 0; it < kTinyIters && rz > 1e-12f * rz0; ++it) {
      const M pw = __shfl(p, (t + 63) & 63, 64), pe = __shfl(p, (t + 1) & 63, 64);
      const M ps = __shfl(p, (t + 64 - nxT) & 63, 64);
      const M pn = __shfl(p, (t + nxT) & 63, 64);
      M ap = dg * p;
      if (iT > 0) ap -= tw * pw;
      if (iT < nxT - 1) ap -= te * pe;
      if (jT > 0) ap -= ts * ps;
      if (jT < nyT - 1) ap -= tn * pn;
      if (!actT) ap = 0.f;
      const M pap = WaveSumF(p * ap);
      if (!(pap > 0.f)) break;
      const M alpha = rz / pap;
      x += alpha * p;
      rr -= alpha * ap;
      z = tdi * rr;
      const M rzn = WaveSumF(rr * z);
      p = z + (rzn / rz) * p;
      rz = rzn;
    }
    x -= WaveSumF(actT ? x : 0.f) / M(nT);
    if (actT) sxa[cs.off[nlev] + t] = x;
  }
  __syncthreads();
  // ---- up: y = xb + P e (into Buf(-1): xb in place for an odd kSweeps, each
  // thread reading and writing only its own cells), kSweeps sweeps; the
  // level's result in xa
#pragma unroll
  for (int l = kCcLevels - 1; l >= 0; --l) {
    if (l >= nlev) continue;
    const int nx = cs.nx[l], ny = cs.ny[l], n = nx * ny, nxc = cs.nx[l + 1];
    const M* r = sr + cs.off[l];
    const M* e = sxa + cs.off[l + 1];
    M *xa = sxa + cs.off[l], *xb = sxb + cs.off[l];
    M* yb = Buf(xa, xb, -1, false);
#pragma unroll
    for (int m = 0; m < CcSlots(l); ++m) {
      const int c = t + m * kCcThreads;
      if (c >= n) continue;
      const int i = ij[l][m] & 0xffff, j = ij[l][m] >> 16, I = i >> 1, J = j >> 1;
      const int Ix = I + ((i & 1) ? 1 : -1), Jy = J + ((j & 1) ? 1 : -1);
      const M ec = e[I + nxc * J];
      const M ex = wx[l][m] > 0.f ? e[Ix + nxc * J] : ec;
      const M ey = wy[l][m] > 0.f ? e[I + nxc * Jy] : ec;
      const M wd = wx[l][m] * wy[l][m];
      const M ed = wd > 0.f ? e[Ix + nxc * Jy] : ec;
      yb[c] = xb[c] + (ec + (3.f * wx[l][m] * (ex - ec) + 3.f * wy[l][m] * (ey - ec) +
                             wd * (ed - ec)) * 0.0625f);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSweeps; ++k) {
      M* dst = Buf(xa, xb, k, false);
      const M* src = Buf(xa, xb, k - 1, false);
#pragma unroll
      for (int m = 0; m < CcSlots(l); ++m) {
        const int c = t + m * kCcThreads;
        if (c < n)
          dst[c] = src[c] + w.w[k] * di[l][m] *
                                (r[c] - Apply(src, l, m, nx, ny, c, ij[l][m] & 0xffff,
                                              ij[l][m] >> 16));
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int m = 0; m < kCcSlots; ++m) {
    const int c = t + m * kCcThreads;
    if (c < n0) cs.out[c] = sxa[c];
  }
}

inline int Blocks1D(long n) { return int(std::min<long>((n + 255) / 256, 4096)); }

}  // namespace

template <class M>
MgPrecondT<M>::MgPrecondT(int nx, int ny, M* kx0, M* ky0, M* dinv0) {
  int cx = nx, cy = ny;
  while (true) {
    MgLevelT<M> L;
    L.nx = cx;
    L.ny = cy;
    L.tg = MakeTileGrid(cx, cy);
    L.mid = MakeTileGrid(cx, cy, kMidY);
    L.tall = MakeTileGrid(cx, cy, kTallY);
    long n = long(cx) * cy;
    if (lv_.empty()) {
      L.kx = kx0;
      L.ky = ky0;
      L.dinv = dinv0;
    } else {
      HG_CHECK(hipMalloc(&L.kx, size_t(cx + 1) * cy * sizeof(M)));
      HG_CHECK(hipMalloc(&L.ky, size_t(cx) * (cy + 1) * sizeof(M)));
      HG_CHECK(hipMalloc(&L.dinv, n * sizeof(M)));
    }
    HG_CHECK(hipMalloc(&L.r, n * sizeof(M)));
    HG_CHECK(hipMalloc(&L.xa, n * sizeof(M)));
    HG_CHECK(hipMalloc(&L.xb, n * sizeof(M)));
    lv_.push_back(L);
    if ((cx & 1) || (cy & 1) || n <= kTinyMax || cx < 4 || cy < 4) break;
    cx /= 2;
    cy /= 2;
  }
  HG_CHECK(hipMalloc(&never_done_, sizeof(int)));
  HG_CHECK(hipMemset(never_done_, 0, sizeof(int)));
  const long nlast = long(lv_.back().nx) * lv_.back().ny;
  if (nlast > kCoarsestMax)
    coarse_sweeps = -1;  // coarsest too large for one workgroup: relax with grid kernels
  // levels 1.. of at most kInKernelMax cells go to the one-workgroup cycle
  // when the hierarchy ends on a tiny grid
  if (nlast <= kTinyMax && lv_.size() > 1) {
    for (int l = 1; l < int(lv_.size()); ++l)
      if (long(lv_[l].nx) * lv_[l].ny <= kInKernelMax) {
        first_in_kernel_ = l;
        break;
      }
  }
  // (a chain from <= kInKernelMax cells down to <= kTinyMax has at most
  // kCcLevels + 1 levels; anything else stays on the grid kernels)
  if (first_in_kernel_ > 0 && int(lv_.size()) - first_in_kernel_ > kCcLevels + 1)
    first_in_kernel_ = -1;
  if (first_in_kernel_ > 0) {
    long cells = 0;
    for (int l = first_in_kernel_; l < int(lv_.size()); ++l) cells += long(lv_[l].nx) * lv_[l].ny;
    if (cells > kLdsCells) first_in_kernel_ = -1;  // cannot happen: 2048 * 4/3 < kLdsCells
  }
}

template <class M>
MgPrecondT<M>::~MgPrecondT() {
  if (never_done_) (void)hipFree(never_done_);
  for (size_t l = 0; l < lv_.size(); ++l) {
    MgLevelT<M>& L = lv_[l];
    for (void* p : {(void*)L.r, (void*)L.xa, (void*)L.xb})
      if (p) (void)hipFree(p);
    if (l > 0)  // level 0's coefficients are the caller's
      for (void* p : {(void*)L.kx, (void*)L.ky, (void*)L.dinv})
        if (p) (void)hipFree(p);
  }
}

template <class M>
void MgPrecondT<M>::Build(hipStream_t st) {
  // grid kernels down to the first level of at most kChainCells cells, then
  // the remaining (small) levels in one workgroup
  constexpr long kChainCells = 4096;
  size_t l = 1;
  for (; l < lv_.size(); ++l) {
    MgLevelT<M>& F = lv_[l - 1];
    MgLevelT<M>& C = lv_[l];
    if (long(C.nx) * C.ny <= kChainCells && lv_.size() - (l - 1) <= size_t(kMaxChain)) break;
    KCoarsen<M><<<Blocks1D(long(C.nx) * C.ny), 256, 0, st>>>(F.nx, F.kx, F.ky, C.nx, C.ny, C.kx,
                                                          C.ky, C.dinv);
  }
  if (l < lv_.size()) {
    ChainSet<M> cs{};
    cs.count = int(lv_.size() - (l - 1));
    for (int k = 0; k < cs.count; ++k) {
      MgLevelT<M>& L = lv_[l - 1 + k];
      cs.nx[k] = L.nx;
      cs.ny[k] = L.ny;
      cs.kx[k] = L.kx;
      cs.ky[k] = L.ky;
      cs.dinv[k] = L.dinv;
    }
    KCoarsenChain<M><<<1, 1024, 0, st>>>(cs);
  }
}

template <class M>
void MgPrecondT<M>::Cycle(int l, const double* r64, float* z, M* out, hipStream_t st,
                          const int* done, const FcgFusion* f) {
  if (!done) done = never_done_;
  const FcgFusion fz = f ? *f : FcgFusion{};
  const int fused = (f && l == 0) ? 1 : 0;
  MgLevelT<M>& L = lv_[l];
  if (l > 0 && l == first_in_kernel_ && use_coarse_kernel) {
    CoarseSetT<M> cs{};
    cs.count = int(lv_.size()) - l;
    cs.bilinear = bilinear ? 1 : 0;
    int off = 0;
    for (int k = 0; k < cs.count; ++k) {
      const MgLevelT<M>& C = lv_[l + k];
      cs.off[k] = off;
      off += C.nx * C.ny;
      cs.nx[k] = C.nx;
      cs.ny[k] = C.ny;
      cs.kx[k] = C.kx;
      cs.ky[k] = C.ky;
      cs.dinv[k] = C.dinv;
      cs.r[k] = C.r;
      cs.xa[k] = C.xa;
      cs.xb[k] = C.xb;
    }
    cs.out = out;
    KCoarseCycle<M><<<1, 1024, 0, st>>>(cs, sweeps, done);
    return;
  }
  const bool coarsest = (l == int(lv_.size()) - 1);
  const int blocks = L.tg.blocks();
  if (coarsest && coarse_sweeps > 0) {
    // r of the coarsest level is always FP32 (lv_.size() > 1 is guaranteed for
    // fine grids; a single-level hierarchy relaxes the FP64 residual below)
    if (l > 0) {
      KCoarsest<M><<<1, kCoarsestMax, 0, st>>>(L.nx, L.ny, L.kx, L.ky, L.dinv, L.r, omega,
                                                coarse_sweeps, out, done);
      return;
    }
  }
  if (coarsest) {
    // the level's result goes to z (FP32) on a single-level hierarchy
    M* target = out;
    if (l == 0) {
      if constexpr (std::is_same<M, float>::value)
        target = z;
      else
        throw std::runtime_error("MgPrecond64: a single-level hierarchy is not supported");
    }
    if (l == 0)
      KSmoothFirst<double, M><<<blocks, kTileThreads, 0, st>>>(L.nx, L.ny, L.tg, r64, L.dinv,
                                                               omega, L.xa, done);
    else
      KSmoothFirst<M, M><<<blocks, kTileThreads, 0, st>>>(L.nx, L.ny, L.tg, L.r, L.dinv, omega,
                                                          L.xa, done);
    // no coarser level (odd or tiny grids): seven more Jacobi sweeps ending
    // in the target, ping-ponging between xa and xb
    M* cur = L.xa;
    const int nsw = 7;
    for (int s = 0; s < nsw; ++s) {
      M* dst = (s == nsw - 1) ? target : (cur == L.xa ? L.xb : L.xa);
      if (dst == cur) dst = (cur == L.xa) ? L.xb : L.xa;
      if (l == 0)
        KJacobi<double, M><<<blocks, kTileThreads, 0, st>>>(L.nx, L.ny, L.tg, L.kx, L.ky, L.dinv,
                                                            r64, cur, omega, dst, done);
      else
        KJacobi<M, M><<<blocks, kTileThreads, 0, st>>>(L.nx, L.ny, L.tg, L.kx, L.ky, L.dinv, L.r,
                                                       cur, omega, dst, done);
      cur = dst;
    }
    if (cur != target)
      HG_CHECK(hipMemcpyAsync(target, cur, size_t(L.nx) * L.ny * sizeof(M),
                              hipMemcpyDeviceToDevice, st));
    return;
  }
  MgLevelT<M>& C = lv_[l + 1];
  // level 0 (FP64 residual) on 64 x kTallY tiles, the FP32 levels on 64 x kMidY;
  // with the FCG update folded into level 0 its stop test moves to level 1
  const bool up = fz.up_on != 0;
  if (l == 0 && up)
    KDown<M, double, kTallY, kFineSweeps, true><<<L.tall.blocks(), kTileThreads, 0, st>>>(
        L.nx, L.ny, L.tall, L.kx, L.ky, L.dinv, r64, fine_sweeps, L.xb, C.nx, C.ny,
        C.r, done, fz, 0);
  else if (l == 0)
    KDown<M, double, kTallY, kFineSweeps><<<L.tall.blocks(), kTileThreads, 0, st>>>(
        L.nx, L.ny, L.tall, L.kx, L.ky, L.dinv, r64, fine_sweeps, L.xb, C.nx, C.ny,
        C.r, done, fz, fused);
  else
    KDown<M, M, kMidY, kSweeps><<<L.mid.blocks(), kTileThreads, 0, st>>>(L.nx, L.ny, L.mid, L.kx, L.ky, L.dinv,
                                                       L.r, sweeps, L.xb, C.nx,
                                                       C.ny, C.r, done, fz,
                                                       (l == 1 && up) ? 1 : 0);
  // the coarser level's correction goes to its xa: its xb holds the x2 that
  // its up-leg reads on tile halos while other workgroups write their output
  Cycle(l + 1, nullptr, nullptr, C.xa, st, done, (l == 0 && up) ? f : nullptr);
  // the up-leg reads x2 (L.xb) and the coarse correction on halos that other
  // workgroups' tiles cover: its output must be a third buffer
  // (level 0 writes z, a separate FP32 buffer)
  if (l > 0 && (out == L.xb || out == C.xa))
    throw std::logic_error("MgPrecond: up-leg output aliases its input");
#define HG_UP(OT, RT, B, TY, SW, WT, TG, R, OUT)                                         \
  KUp<M, OT, RT, B, TY, SW><<<TG.blocks(), kTileThreads, 0, st>>>(                       \
      L.nx, L.ny, TG, L.kx, L.ky, L.dinv, R, L.xb, C.xa, C.kx, C.ky, C.nx, C.ny, WT, OUT, \
      done, fz, fused)
  if (l == 0) {
    // after the folded update the level's residual is the one it wrote
    const double* r0 = up ? fz.up_r : r64;
    if (bilinear)
      HG_UP(float, double, true, kTallY, kFineSweeps, fine_sweeps, L.tall, r0, z);
    else
      HG_UP(float, double, false, kTallY, kFineSweeps, fine_sweeps, L.tall, r0, z);
  } else {
    if (bilinear)
      HG_UP(M, M, true, kMidY, kSweeps, sweeps, L.mid, L.r, out);
    else
      HG_UP(M, M, false, kMidY, kSweeps, sweeps, L.mid, L.r, out);
  }
#undef HG_UP
}

template <class M>
void MgPrecondT<M>::Apply(const double* r, float* z, hipStream_t st, const int* d_done) {
  Cycle(0, r, z, nullptr, st, d_done);
}

template <class M>
bool MgPrecondT<M>::ApplyFcg(const double* r, float* z, hipStream_t st, const int* d_done,
                             const FcgFusion& f, int* ndir) {
  if (lv_.size() < 2) return false;  // no KDown / KUp pair on the finest level
  Cycle(0, r, z, nullptr, st, d_done, &f);
  *ndir = lv_[0].tall.blocks();
  return true;
}

template <class M>
double MgPrecondT<M>::TimeKernel(const char* name, int reps, const double* r, hipStream_t st) {
  if (lv_.size() < 2) throw std::runtime_error("TimeKernel: single-level hierarchy");
  const std::string k(name);
  MgLevelT<M>& L = lv_[0];
  MgLevelT<M>& C = lv_[1];
  auto launch = [&]() {
    if (k == "KDown<double>") {
      KDown<M, double, kTallY, kFineSweeps><<<L.tall.blocks(), kTileThreads, 0, st>>>(
          L.nx, L.ny, L.tall, L.kx, L.ky, L.dinv, r, fine_sweeps, L.xb, C.nx, C.ny,
          C.r, never_done_, FcgFusion{}, 0);
    } else if (k == "KUp<double, true>") {
      KUp<M, M, double, true, kTallY, kFineSweeps><<<L.tall.blocks(), kTileThreads, 0, st>>>(
          L.nx, L.ny, L.tall, L.kx, L.ky, L.dinv, r, L.xb, C.xa, C.kx, C.ky, C.nx, C.ny,
          fine_sweeps, L.xa, never_done_, FcgFusion{}, 0);
    } else if (k == "KCoarseCycle" && first_in_kernel_ > 0) {
      Cycle(first_in_kernel_, nullptr, nullptr, lv_[first_in_kernel_].xa, st, nullptr);
    } else {
      throw std::runtime_error("TimeKernel: unknown kernel " + k);
    }
  };
  launch();  // warm-up
  hipEvent_t e0, e1;
  HG_CHECK(hipEventCreate(&e0));
  HG_CHECK(hipEventCreate(&e1));
  HG_CHECK(hipEventRecord(e0, st));
  for (int i = 0; i < reps; ++i) launch();
  HG_CHECK(hipEventRecord(e1, st));
  HG_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  HG_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return double(ms) / reps;
}

template class MgPrecondT<float>;
template class MgPrecondT<double>;

}  // namespace hg
