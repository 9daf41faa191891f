// step_kernels.hip -- kernels of the hydro2d time step.
//
// One kernel per stage of FluidSimple::MakeIteration (fluid.hpp:814-1158),
// AdvectionSolverMultiExplicit::MakeIteration (advection.hpp:440-539) and
// hydro::UpdateFluidProperties (hydro2d.hpp:1404-1430).  Every value is
// computed with the reference's operation order (this file is compiled with
// -ffp-contract=off), so on power-of-two meshes the FP64 results equal the
// reference's bit for bit.  Fields are SoA, x fastest; faces are the
// x-faces followed by the y-faces.
//
// The state (velocity, pressure, volume flux, partial densities, properties)
// is FP64.  The intermediates of a SIMPLE iteration -- per-step face fields,
// cell gradients, momentum rows -- have a storage type T: double in the FP64
// path, float in the mixed path.  Arithmetic that cancels (differences of
// state values, residuals, corrections) stays FP64 in both; the red-black
// velocity sweep and the advection limiter compute in T.

#include "kernels.h"

namespace hg {

namespace {

constexpr int kBlock = 256;

inline dim3 Grid(long n) { return dim3(unsigned((n + kBlock - 1) / kBlock)); }

// Velocity component `comp` of the fluid condition on side s at position q
// (j on the left / right, i on the bottom / top).
__device__ __forceinline__ double BVel(const Bc& bc, const Geo& g, int comp, int s, int q) {
  return bc.fvel[comp] ? bc.fvel[comp][BoundaryIndex(g, s, q)] : bc.vel[comp][s];
}

// ---------------------------------------------------------------------------
// Face interpolation of the per-step inputs (fluid.hpp:602-641).
template <class T>
__global__ void __launch_bounds__(kBlock)
KForceFaces(Geo g, const double* __restrict__ fx, const double* __restrict__ fy,
            const double* __restrict__ mu, T* __restrict__ ffex, T* __restrict__ ffey,
            T* __restrict__ ffmu) {
  const int nf = g.nfx + g.nx * (g.ny + 1);
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int cm, cp;
  if (f < g.nfx) {
    int i = f % (g.nx + 1), j = f / (g.nx + 1);
    cm = (i > 0) ? (i - 1) + g.nx * j : -1;
    cp = (i < g.nx) ? i + g.nx * j : -1;
  } else {
    int r = f - g.nfx;
    int i = r % g.nx, j = r / g.nx;
    cm = (j > 0) ? i + g.nx * (j - 1) : -1;
    cp = (j < g.ny) ? i + g.nx * j : -1;
  }
  if (cm >= 0 && cp >= 0) {
    ffex[f] = T(Mid(fx[cm], fx[cp]));
    ffey[f] = T(Mid(fy[cm], fy[cp]));
    ffmu[f] = T(Mid(mu[cm], mu[cp]));
  } else {
    int c = cm >= 0 ? cm : cp;  // zero-derivative condition: cell value
    ffex[f] = T(fx[c]);
    ffey[f] = T(fy[c]);
    ffmu[f] = T(mu[c]);
  }
}

// Restored force (fluid.hpp:612-626): sum_f S_f (f_f . n_f) |x_c - x_f| / V.
template <class T>
__global__ void __launch_bounds__(kBlock)
KRestoredForce(Geo g, const T* __restrict__ ffex, const T* __restrict__ ffey,
               T* __restrict__ frx, T* __restrict__ fry) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nx * g.ny) return;
  const int i = c % g.nx, j = c / g.nx;
  const int fw = XFace(g, i, j), fe = XFace(g, i + 1, j);
  const int fs = YFace(g, i, j), fn = YFace(g, i, j + 1);
  // x-faces: S = (sx, -0), n = (1, -0): S*(f.n) has only an x component.
  double sx = 0., sy = 0.;
  sx += g.sx * double(ffex[fw]) * g.rdist_x;
  sx += g.sx * double(ffex[fe]) * g.rdist_x;
  sy += g.sy * double(ffey[fs]) * g.rdist_y;
  sy += g.sy * double(ffey[fn]) * g.rdist_y;
  frx[c] = T(OverV(g, sx));
  fry[c] = T(OverV(g, sy));
}

// Face interpolation with zero-derivative condition for a pair of cell fields.
template <class T>
__global__ void __launch_bounds__(kBlock)
KFacesDeriv0Pair(Geo g, const T* __restrict__ ax, const T* __restrict__ ay,
                 T* __restrict__ fx, T* __restrict__ fy) {
  const int nf = g.nfx + g.nx * (g.ny + 1);
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int cm, cp;
  if (f < g.nfx) {
    int i = f % (g.nx + 1), j = f / (g.nx + 1);
    cm = (i > 0) ? (i - 1) + g.nx * j : -1;
    cp = (i < g.nx) ? i + g.nx * j : -1;
  } else {
    int r = f - g.nfx;
    int i = r % g.nx, j = r / g.nx;
    cm = (j > 0) ? i + g.nx * (j - 1) : -1;
    cp = (j < g.ny) ? i + g.nx * j : -1;
  }
  if (cm >= 0 && cp >= 0) {
    fx[f] = T(Mid(double(ax[cm]), double(ax[cp])));
    fy[f] = T(Mid(double(ay[cm]), double(ay[cp])));
  } else {
    int c = cm >= 0 ? cm : cp;
    fx[f] = ax[c];
    fy[f] = ay[c];
  }
}

// Pressure gradient (extrapolation condition) and velocity gradients (value
// condition) at cells (fluid.hpp:827-832, 840-842; conv_diff.hpp:135).
// FV: the condition has per-face velocities (inlets or outlets); without
// them the side's constant velocity needs no per-cell load.
// The differences of the FP64 fields are taken in FP64; the gradients are
// stored as T.
template <bool FV, class T>
__global__ void __launch_bounds__(kBlock)
KCellGrads(Geo g, Bc bc, const double* __restrict__ p, const double* __restrict__ u0,
           const double* __restrict__ u1, T* __restrict__ gpx, T* __restrict__ gpy,
           T* __restrict__ g0x, T* __restrict__ g0y, T* __restrict__ g1x,
           T* __restrict__ g1y) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nx * g.ny) return;
  const int i = c % g.nx, j = c / g.nx;
  // all three fields' neighbourhoods are loaded before any is used
  const Cross xp = LoadCross(p, g.nx, g.ny, i, j);
  const Cross x0 = LoadCross(u0, g.nx, g.ny, i, j);
  const Cross x1 = LoadCross(u1, g.nx, g.ny, i, j);
  double f[4], gx, gy;
  CellFacesV<2>(g, xp, i, j, nullptr, f);
  GradFromFaces(g, f, gx, gy);
  gpx[c] = T(gx);
  gpy[c] = T(gy);
  double bv0[4], bv1[4];
  const int qs[4] = {j, j, i, i};
  for (int s = 0; s < 4; ++s) {
    bv0[s] = FV ? BVel(bc, g, 0, s, qs[s]) : bc.vel[0][s];
    bv1[s] = FV ? BVel(bc, g, 1, s, qs[s]) : bc.vel[1][s];
  }
  CellFacesV<0>(g, x0, i, j, bv0, f);
  GradFromFaces(g, f, gx, gy);
  g0x[c] = T(gx);
  g0y[c] = T(gy);
  CellFacesV<0>(g, x1, i, j, bv1, f);
  GradFromFaces(g, f, gx, gy);
  g1x[c] = T(gx);
  g1y[c] = T(gy);
}

template <class T>
struct MomIn {
  const double *u0, *u1;
  const T *g0x, *g0y, *g1x, *g1y, *gpx, *gpy, *frx, *fry, *ffmu;
  const double *flux, *rho, *utp0, *utp1, *utc0, *utc1;
  double bdf0, bdf1, bdf2, relax;
};
template <class T>
struct MomOut {
  T *s, *w, *c, *e, *n, *k0, *k1;
  double* ap;  // optional: a_P of the relaxed rows
};

// Explicit viscosity + force assembly + momentum rows (fluid.hpp:835-874,
// conv_diff.hpp:130-243).  Row coefficients are shared by both velocity
// components (their BCs, fluxes and densities coincide); constants differ.
// The row constants -- the delta form's residual at the previous iterate --
// are formed in FP64 from the FP64 coefficients and stored as T with them.
template <class T>
__global__ void __launch_bounds__(kBlock)
KMomentum(Geo g, Bc bc, MomIn<T> in, MomOut<T> out) {
  const int nx = g.nx, ny = g.ny;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  const int i = c % nx, j = c / nx;
  const bool hw = i > 0, he = i < nx - 1, hs = j > 0, hn = j < ny - 1;
  const int cw = c - 1, ce = c + 1, cs = c - nx, cn = c + nx;
  const int fw = XFace(g, i, j), fe = XFace(g, i + 1, j);
  const int fs = YFace(g, i, j), fn = YFace(g, i, j + 1);
  const double V = g.V;
  // every input of the cell is loaded here, neighbours at clamped indices
  // (a neighbour value is used only where the neighbour exists), so the
  // thread waits on memory once; below, values are selected, not loaded
  const int cwc = hw ? cw : c, cec = he ? ce : c, csc = hs ? cs : c, cnc = hn ? cn : c;
  const double muw = in.ffmu[fw], mue = in.ffmu[fe], mus = in.ffmu[fs], mun = in.ffmu[fn];
  const double Fw = in.flux[fw], Fe = in.flux[fe], Fs = in.flux[fs], Fn = in.flux[fn];
  // gradients: g0x at C, W, E; g0y and g1x at C, W, E, S, N; g1y at C, S, N
  const double g0xC = in.g0x[c], g0xW = in.g0x[cwc], g0xE = in.g0x[cec];
  const double g0yC = in.g0y[c], g0yW = in.g0y[cwc], g0yE = in.g0y[cec], g0yS = in.g0y[csc],
               g0yN = in.g0y[cnc];
  const double g1xC = in.g1x[c], g1xW = in.g1x[cwc], g1xE = in.g1x[cec], g1xS = in.g1x[csc],
               g1xN = in.g1x[cnc];
  const double g1yC = in.g1y[c], g1yS = in.g1y[csc], g1yN = in.g1y[cnc];
  const double u0C = in.u0[c], u0W = in.u0[cwc], u0E = in.u0[cec], u0S = in.u0[csc],
               u0N = in.u0[cnc];
  const double u1C = in.u1[c], u1W = in.u1[cwc], u1E = in.u1[cec], u1S = in.u1[csc],
               u1N = in.u1[cnc];
  const double gpxc = in.gpx[c], gpyc = in.gpy[c], frxc = in.frx[c], fryc = in.fry[c];
  const double rhoc = in.rho[c];
  const double utp0c = in.utp0[c], utp1c = in.utp1[c], utc0c = in.utc0[c], utc1c = in.utc1[c];
  // face value of a cell gradient with the zero-derivative condition
  // (fluid.hpp:842): Mid(lower-index cell, higher-index cell) on inner faces,
  // the cell's own value on boundary faces
  auto FG = [](double vc, double vnb, bool nb_lower, bool inner) {
    return inner ? Mid(nb_lower ? vnb : vc, nb_lower ? vc : vnb) : vc;
  };

  // --- explicit viscosity (transpose stress), component by component
  double fox = 0., foy = 0.;
  {
    // d = 0: weights mu_f * S_out[0] on the W/E faces.
    double ww = muw * (-g.sx), we = mue * g.sx;
    double gwx = FG(g0xC, g0xW, true, hw), gex = FG(g0xC, g0xE, false, he);
    double gwy = FG(g0yC, g0yW, true, hw), gey = FG(g0yC, g0yE, false, he);
    double sx = gwx * ww;
    sx += gex * we;
    double sy = gwy * ww;
    sy += gey * we;
    fox += OverV(g, sx);
    foy += OverV(g, sy);
    // d = 1: weights mu_f * S_out[1] on the S/N faces.
    double ws = mus * (-g.sy), wn = mun * g.sy;
    double gsx = FG(g1xC, g1xS, true, hs), gnx = FG(g1xC, g1xN, false, hn);
    double gsy = FG(g1yC, g1yS, true, hs), gny = FG(g1yC, g1yN, false, hn);
    sx = gsx * ws;
    sx += gnx * wn;
    sy = gsy * ws;
    sy += gny * wn;
    fox += OverV(g, sx);
    foy += OverV(g, sy);
  }
  // --- force assembly: += -grad p + restored force (+ zero surface tension
  //     and zero volume-source compensation on this path)
  fox += gpxc * (-1.) + frxc;
  foy += gpyc * (-1.) + fryc;

  // --- face expressions.  Convective (upwind deferred, threshold 1e-10,
  //     solver.hpp:201-243) and diffusive (solver.hpp:283-341).
  const bool inner[4] = {hw, he, hs, hn};
  const double Ff[4] = {Fw, Fe, Fs, Fn};
  // the cell's and the neighbour's values per face (W, E: x-faces; S, N)
  const double a0nb[4] = {g0xW, g0xE, g0yS, g0yN}, a0c[4] = {g0xC, g0xC, g0yC, g0yC};
  const double a1nb[4] = {g1xW, g1xE, g1yS, g1yN}, a1c[4] = {g1xC, g1xC, g1yC, g1yC};
  const double mu[4] = {muw, mue, mus, mun};
  const double outf[4] = {-1., 1., -1., 1.};
  const int side[4] = {kLeft, kRight, kBottom, kTop};
  double Cc = 0., Dc = 0., Cnb[4] = {0, 0, 0, 0}, Dnb[4] = {0, 0, 0, 0};
  double Ck0 = 0., Ck1 = 0., Dk0 = 0., Dk1 = 0.;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const bool xface = k < 2;
    const double A = xface ? g.ax : g.ay;
    const double F = Ff[k];
    const double o = outf[k];
    const bool self_m = (k == 1 || k == 3);
    if (inner[k]) {
      // gradient values at cell m (minus side) and p (plus side) of the face
      const double g0m = self_m ? a0c[k] : a0nb[k], g0p = self_m ? a0nb[k] : a0c[k];
      const double g1m = self_m ? a1c[k] : a1nb[k], g1p = self_m ? a1nb[k] : a1c[k];
      double am, ap, ak0, ak1;
      if (F > 1e-10) {
        am = 1.;
        ap = 0.;
        ak0 = -(xface ? g0m * g.tx0 : g0m * g.ty0);
        ak1 = -(xface ? g1m * g.tx0 : g1m * g.ty0);
      } else if (F < -1e-10) {
        am = 0.;
        ap = 1.;
        ak0 = -(xface ? g0p * g.tx1 : g0p * g.ty1);
        ak1 = -(xface ? g1p * g.tx1 : g1p * g.ty1);
      } else {
        am = 0.5;
        ap = 0.5;
        ak0 = 0.;
        ak1 = 0.;
      }
      const double cam = am * F, cap = ap * F;
      const double al = 1. / (xface ? g.dnx : g.dny);
      const double dcm = ((-al) * (-mu[k])) * A;
      const double dcp = (al * (-mu[k])) * A;
      const double dck = (0. * (-mu[k])) * A;
      Cc += (self_m ? cam : cap) * o;
      Cnb[k] = (self_m ? cap : cam) * o;
      Dc += (self_m ? dcm : dcp) * o;
      Dnb[k] = (self_m ? dcp : dcm) * o;
      Ck0 += (ak0 * F) * o;
      Ck1 += (ak1 * F) * o;
      Dk0 += dck * o;
      Dk1 += dck * o;
    } else {
      const int sd = side[k];
      const double al = g.bal[sd];
      const double v0 = BVel(bc, g, 0, sd, sd < 2 ? j : i), v1 = BVel(bc, g, 1, sd, sd < 2 ? j : i);
      Ck0 += (v0 * F) * o;
      Ck1 += (v1 * F) * o;
      Dc += (((-al) * (-mu[k])) * A) * o;
      Dk0 += (((al * v0) * (-mu[k])) * A) * o;
      Dk1 += (((al * v1) * (-mu[k])) * A) * o;
    }
  }
  const double rho = rhoc;
  const double cen = OverV(g, Dc) + ((OverV(g, Cc) + in.bdf2) * rho);
  double nb[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) nb[k] = OverV(g, Dnb[k]) + (OverV(g, Cnb[k]) * rho);
  // constants, then the delta form (row evaluated at the previous iterate in
  // the order S, W, C, E, N) and under-relaxation of the diagonal
  const double un0 = in.bdf0 * utp0c + in.bdf1 * utc0c;
  const double un1 = in.bdf0 * utp1c + in.bdf1 * utc1c;
  double k0 = (OverV(g, Dk0) + ((OverV(g, Ck0) + un0) * rho)) + (-fox);
  double k1 = (OverV(g, Dk1) + ((OverV(g, Ck1) + un1) * rho)) + (-foy);
  if (hs) { k0 += u0S * nb[2]; k1 += u1S * nb[2]; }
  if (hw) { k0 += u0W * nb[0]; k1 += u1W * nb[0]; }
  k0 += u0C * cen;
  k1 += u1C * cen;
  if (he) { k0 += u0E * nb[1]; k1 += u1E * nb[1]; }
  if (hn) { k0 += u0N * nb[3]; k1 += u1N * nb[3]; }
  const double rs = hs ? nb[2] : 0., rw = hw ? nb[0] : 0., rc = cen / in.relax;
  const double re = he ? nb[1] : 0., rn = hn ? nb[3] : 0.;
  out.s[c] = T(rs);
  out.w[c] = T(rw);
  out.c[c] = T(rc);
  out.e[c] = T(re);
  out.n[c] = T(rn);
  out.k0[c] = T(k0);
  out.k1[c] = T(k1);
  if (out.ap) {
    // a_P = CoeffSum of the relaxed row, averaged over the two velocity
    // components (identical rows): fluid.hpp:876-883
    double cs = 0.;
    if (hs) cs += rs;
    if (hw) cs += rw;
    cs += rc;
    if (he) cs += re;
    if (hn) cs += rn;
    double sum = 0.;
    sum += cs;
    sum += cs;
    out.ap[c] = sum / 2.;
  }
}

// Diagonal coefficient a_P = mean of the rows' CoeffSum (fluid.hpp:876-884).
template <class T>
__global__ void __launch_bounds__(kBlock)
KRhieChow(Geo g, Bc bc, const double* __restrict__ u0, const double* __restrict__ u1,
          const T* __restrict__ gpx, const T* __restrict__ gpy,
          const T* __restrict__ ffrx, const T* __restrict__ ffry,
          const T* __restrict__ ffex, const T* __restrict__ ffey,
          const double* __restrict__ p, const double* __restrict__ ap, double rc, double mvx,
          double mvy, double* __restrict__ fstar, double* __restrict__ fcoef) {
  const int nf = g.nfx + g.nx * (g.ny + 1);
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const bool xface = f < g.nfx;
  int cm, cp, sd, q;
  if (xface) {
    int i = f % (g.nx + 1), j = f / (g.nx + 1);
    cm = (i > 0) ? (i - 1) + g.nx * j : -1;
    cp = (i < g.nx) ? i + g.nx * j : -1;
    sd = (i == 0) ? kLeft : kRight;
    q = j;
  } else {
    int r = f - g.nfx;
    int i = r % g.nx, j = r / g.nx;
    cm = (j > 0) ? i + g.nx * (j - 1) : -1;
    cp = (j < g.ny) ? i + g.nx * j : -1;
    sd = (j == 0) ? kBottom : kTop;
    q = i;
  }
  const double S = xface ? g.sx : g.sy;
  double fs, fc = 0.;
  if (cm >= 0 && cp >= 0) {
    // u*_f . S with u*_f = (u[cm] + u[cp]) / 2 component-wise
    const double ufn = xface ? Mid(u0[cm], u0[cp]) : Mid(u1[cm], u1[cp]);
    const double vfi = ufn * S;
    const double gpf = xface ? Mid(double(gpx[cm]), double(gpx[cp]))
                             : Mid(double(gpy[cm]), double(gpy[cp]));
    const double frf = xface ? double(ffrx[f]) : double(ffry[f]);
    const double wide = (gpf - frf) * S;
    const double dn = xface ? g.dnx : g.dny;
    const double A = xface ? g.ax : g.ay;
    const double fef = xface ? double(ffex[f]) : double(ffey[f]);
    const double compact = (p[cp] - p[cm]) / dn * A - fef * S;
    const double af = Mid(ap[cm], ap[cp]);
    fs = vfi + rc * (wide - compact) / af + 0;
    fc = -A / (dn * af);
  } else {
    const double vb = BVel(bc, g, xface ? 0 : 1, sd, q);
    fs = vb * S;
  }
  fs -= (xface ? mvx : mvy) * S;
  fstar[f] = fs;
  fcoef[f] = fc;
}

// Pressure-correction rows (fluid.hpp:972-1027) with an optional fixed
// pressure cell (Expression::SetKnownValue, linear.hpp:237-249).
__global__ void __launch_bounds__(kBlock)
KPressureRows(Geo g, const double* __restrict__ fstar, const double* __restrict__ fcoef,
              int pinned, double pval, double* __restrict__ s, double* __restrict__ w,
              double* __restrict__ cc, double* __restrict__ e, double* __restrict__ n,
              double* __restrict__ k) {
  const int nx = g.nx, ny = g.ny;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  const int i = c % nx, j = c / nx;
  const int fw = XFace(g, i, j), fe = XFace(g, i + 1, j);
  const int fs = YFace(g, i, j), fn = YFace(g, i, j + 1);
  double cen = 0., kon = 0.;
  double nw = 0., ne = 0., ns = 0., nn = 0.;
  if (i > 0) { cen += fcoef[fw] * -1.; nw = (-fcoef[fw]) * -1.; }
  kon += fstar[fw] * -1.;
  if (i < nx - 1) { cen += (-fcoef[fe]) * 1.; ne = fcoef[fe] * 1.; }
  kon += fstar[fe] * 1.;
  if (j > 0) { cen += fcoef[fs] * -1.; ns = (-fcoef[fs]) * -1.; }
  kon += fstar[fs] * -1.;
  if (j < ny - 1) { cen += (-fcoef[fn]) * 1.; nn = fcoef[fn] * 1.; }
  kon += fstar[fn] * 1.;
  kon += -(0. * g.V);  // volume source is zero on this path
  if (pinned >= 0) {
    if (c == pinned) {
      nw = ne = ns = nn = 0.;
      cen = 1.;
      kon = -pval;
    } else {
      if (c + nx == pinned) { kon += pval * nn; nn = 0.; }
      if (c + 1 == pinned && i < nx - 1) { kon += pval * ne; ne = 0.; }
      if (c == pinned + 1 && i > 0) { kon += pval * nw; nw = 0.; }
      if (c == pinned + nx) { kon += pval * ns; ns = 0.; }
    }
  }
  s[c] = ns;
  w[c] = nw;
  cc[c] = cen;
  e[c] = ne;
  n[c] = nn;
  k[c] = kon;
}

// Correction (fluid.hpp:1033-1057) for cells, with the convergence
// indicator max_c |u_curr - u_prev| (solver.hpp:804-813): per-block maxima,
// folded by KMaxFold (max is order-independent, so the result is exact).
// The constant of the pressure rows alone (KPressureRows' k, no pinned
// cell): what the CG solvers need, which take the operator from the face
// coefficients.
__global__ void __launch_bounds__(kBlock)
KPressureRhs(Geo g, const double* __restrict__ fstar, double* __restrict__ k, int negate) {
  const int nx = g.nx, ny = g.ny;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  const int i = c % nx, j = c / nx;
  double kon = 0.;
  kon += fstar[XFace(g, i, j)] * -1.;
  kon += fstar[XFace(g, i + 1, j)] * 1.;
  kon += fstar[YFace(g, i, j)] * -1.;
  kon += fstar[YFace(g, i, j + 1)] * 1.;
  kon += -(0. * g.V);
  k[c] = negate ? -kon : kon;
}

__global__ void __launch_bounds__(kBlock)
KCorrectCells(Geo g, const double* __restrict__ pc, const double* __restrict__ ap,
              const double* __restrict__ pprev, double alpha_p, double* __restrict__ pcurr,
              double* __restrict__ u0, double* __restrict__ u1, const double* __restrict__ up0,
              const double* __restrict__ up1, double* __restrict__ part) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double d = 0.;
  if (c < g.nx * g.ny) {
    const int i = c % g.nx, j = c / g.nx;
    pcurr[c] = pprev[c] + alpha_p * pc[c];
    double f[4], gx, gy;
    CellFaces<2>(g, pc, i, j, nullptr, f);
    GradFromFaces(g, f, gx, gy);
    const double a = -ap[c];
    const double n0 = u0[c] + gx / a;
    const double n1 = u1[c] + gy / a;
    u0[c] = n0;
    u1[c] = n1;
    const double d0 = up0[c] - n0, d1 = up1[c] - n1;
    double s2 = 0.;
    s2 += d0 * d0;
    s2 += d1 * d1;
    d = sqrt(s2);
    if (!(d == d)) d = 0.;  // std::max(res, NaN) keeps res
  }
  d = WaveMax(d);
  __shared__ double red[kBlock / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < kBlock / 64; ++w) m = fmax(m, red[w]);
    part[blockIdx.x] = m;
  }
}

// Folds the per-block maxima of KCorrectCells into rs (one workgroup of
// kFoldThreads: the partials number N / 256, and a wider workgroup folds
// them sooner; a maximum does not depend on the order).
constexpr int kFoldThreads = 1024;
__global__ void __launch_bounds__(kFoldThreads)
KMaxFold(const double* __restrict__ part, int n, double* __restrict__ rs) {
  // the first eight partials of the thread are loaded at once (indices
  // clamped into the array, selected afterwards), the rest in a loop
  constexpr int K = 8;
  const int t = threadIdx.x, last = n > 0 ? n - 1 : 0;
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = part[min(t + k * kFoldThreads, last)];
  double m = 0.;
#pragma unroll
  for (int k = 0; k < K; ++k)
    if (t + k * kFoldThreads < n) m = fmax(m, v[k]);
  for (int b = t + K * kFoldThreads; b < n; b += kFoldThreads) m = fmax(m, part[b]);
  m = WaveMax(m);
  __shared__ double red[kFoldThreads / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0];
    for (int w = 1; w < kFoldThreads / 64; ++w) t = fmax(t, red[w]);
    *rs = t;
  }
}

// Divergence-free volume flux: the flux row evaluated at p' (fluid.hpp:1053-1056).
// flux32 (optional): an FP32 copy, the advection's input in the mixed path.
__global__ void __launch_bounds__(kBlock)
KCorrectFaces(Geo g, const double* __restrict__ fstar, const double* __restrict__ fcoef,
              const double* __restrict__ pc, double* __restrict__ flux,
              float* __restrict__ flux32) {
  const int nf = g.nfx + g.nx * (g.ny + 1);
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int cm, cp;
  if (f < g.nfx) {
    int i = f % (g.nx + 1), j = f / (g.nx + 1);
    cm = (i > 0) ? (i - 1) + g.nx * j : -1;
    cp = (i < g.nx) ? i + g.nx * j : -1;
  } else {
    int r = f - g.nfx;
    int i = r % g.nx, j = r / g.nx;
    cm = (j > 0) ? i + g.nx * (j - 1) : -1;
    cp = (j < g.ny) ? i + g.nx * j : -1;
  }
  double v = fstar[f];
  if (cm >= 0 && cp >= 0) {
    v += pc[cm] * (-fcoef[f]);
    v += pc[cp] * fcoef[f];
  }
  flux[f] = v;
  if (flux32) flux32[f] = float(v);
}

// Initial volume flux u_f . S - meshvel . S (fluid.hpp:770-784).
__global__ void __launch_bounds__(kBlock)
KInitialFlux(Geo g, Bc bc, const double* __restrict__ u0, const double* __restrict__ u1,
             double mvx, double mvy, double* __restrict__ flux) {
  const int nf = g.nfx + g.nx * (g.ny + 1);
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const bool xface = f < g.nfx;
  int cm, cp, sd, q;
  if (xface) {
    int i = f % (g.nx + 1), j = f / (g.nx + 1);
    cm = (i > 0) ? (i - 1) + g.nx * j : -1;
    cp = (i < g.nx) ? i + g.nx * j : -1;
    sd = (i == 0) ? kLeft : kRight;
    q = j;
  } else {
    int r = f - g.nfx;
    int i = r % g.nx, j = r / g.nx;
    cm = (j > 0) ? i + g.nx * (j - 1) : -1;
    cp = (j < g.ny) ? i + g.nx * j : -1;
    sd = (j == 0) ? kBottom : kTop;
    q = i;
  }
  const double S = xface ? g.sx : g.sy;
  double un;
  if (cm >= 0 && cp >= 0)
    un = xface ? Mid(u0[cm], u0[cp]) : Mid(u1[cm], u1[cp]);
  else
    un = BVel(bc, g, xface ? 0 : 1, sd, q);
  double v = un * S;
  v -= (xface ? mvx : mvy) * S;
  flux[f] = v;
}

// ---------------------------------------------------------------------------
// Advection (advection.hpp:440-478): gradient of the zero-derivative face
// interpolation, then Superbee face values and the explicit update.

template <class R>
__device__ __forceinline__ R Superbee(R p, R q) {
  const R two = R(2), zero = R(0);
  if (p > zero && q > zero) return fmax(fmin(two * p, q), fmin(p, two * q));
  if (p < zero && q < zero) return -fmax(fmin(-two * p, -q), fmin(-p, -two * q));
  return zero;
}

// Superbee face value between P (cell 0) and E (cell 1) (solver.hpp:576-592).
template <class R>
__device__ __forceinline__ R SuperbeeFace(R uP, R uE, R gPn, R gEn, R r0, R r1, R vf) {
  const R half = R(0.5), four = R(4), eps = R(1e-8);
  if (vf > eps) return uP + half * Superbee<R>(uE - uP, -four * (gPn * r0) - (uE - uP));
  if (vf < -eps) return uE - half * Superbee<R>(uE - uP, four * (gEn * r1) - (uE - uP));
  return half * (uP + uE);
}

template <class R>
__device__ __forceinline__ R MidR(R um, R up) { return um * R(0.5) + up * R(0.5); }
// x / V in R, as a product when that is exact (V a power of two)
template <class R>
__device__ __forceinline__ R OverVR(const Geo& g, R x) {
  return g.pow2V ? x * R(g.invV) : x / R(g.V);
}

struct AdvFields {
  const double* u[3];
  double* un[3];
  // inlet values per boundary face of each field (nullptr: no inlet) and
  // which sides are inlets; other boundary faces have zero derivative
  const double* bval[3];
  int inlet[4];
};

// One advection sub-step (advection.hpp:440-539) of NF fields sharing the
// volume flux.  The cell gradients with the zero-derivative condition
// (CellFaces<1> + GradFromFaces) are formed on chip instead of being written
// to HBM and read back: the block loads u on its 64 x kAdvTY tile with a
// two-cell halo into LDS (every field at once, one barrier); each thread then
// owns a vertical segment of kAdvSeg cells of one column and forms, from its
// column's values and its rows' two neighbours either side, the y-gradients
// of the segment plus one cell each end and the x-gradients of each cell and
// its west and east neighbours, in registers.  Every value is computed with
// the same operations as the unfused form, so the result is bit for bit the
// same; a gradient is formed by each cell that needs it (about three times
// over), which costs less than a pass of LDS stores, a barrier and LDS
// reloads (the kernel's limit is LDS traffic, not arithmetic).
// R is the arithmetic type of the gradients and the Superbee face values,
// and the type of the flux read (float in the mixed path).  The fields stay
// FP64 and the update sums the four face terms in FP64: a face term is the
// same number for the two cells it separates, so the fields' volume sums
// are conserved to FP64 rounding in both paths.
constexpr int kAdvTX = 64, kAdvTY = 8;
constexpr int kAdvSeg = kAdvTY / (kBlock / kAdvTX);  // cells per thread, one column
static_assert(kBlock % kAdvTX == 0 && kAdvTY % (kBlock / kAdvTX) == 0, "advection tile");

template <int NF, bool IN, class R>
__global__ void __launch_bounds__(kBlock)
KAdvStep(Geo g, AdvFields fl, const R* __restrict__ flux, double dt) {
  constexpr int UX = kAdvTX + 4, UY = kAdvTY + 4, S = kAdvSeg;
  __shared__ double su[NF][UY][UX];
  const int nx = g.nx, ny = g.ny;
  const int ntx = (nx + kAdvTX - 1) / kAdvTX;
  const int i0 = (blockIdx.x % ntx) * kAdvTX, j0 = (blockIdx.x / ntx) * kAdvTY;
  // the thread's cells: column li, tile rows lj0 .. lj0 + S - 1
  const int li = threadIdx.x % kAdvTX, lj0 = (threadIdx.x / kAdvTX) * S;
  const int i = i0 + li;
  // every global load of the block is issued here, before any value is used:
  // each field on the halo-2 tile and the face fluxes of the thread's cells
  // (so the block waits on memory once, not once per element)
  constexpr int NU = UX * UY, KU = (NU + kBlock - 1) / kBlock;
  double ureg[NF][KU];
#pragma unroll
  for (int q = 0; q < NF; ++q)
#pragma unroll
    for (int k = 0; k < KU; ++k) {
      const int t = threadIdx.x + k * kBlock;
      const int ii = i0 + t % UX - 2, jj = j0 + t / UX - 2;
      ureg[q][k] = (t < NU && ii >= 0 && ii < nx && jj >= 0 && jj < ny) ? fl.u[q][ii + nx * jj]
                                                                         : 0.;
    }
  // x-faces west and east of each cell; the segment's S + 1 y-faces
  R vw_[S], ve_[S], vy_[S + 1];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int j = j0 + lj0 + s;
    const bool in = i < nx && j < ny;
    vw_[s] = in ? flux[XFace(g, i, j)] : R(0);
    ve_[s] = in ? flux[XFace(g, i + 1, j)] : R(0);
  }
#pragma unroll
  for (int s = 0; s <= S; ++s) {
    const int j = j0 + lj0 + s;
    vy_[s] = (i < nx && j <= ny) ? flux[YFace(g, i, j)] : R(0);
  }
#pragma unroll
  for (int q = 0; q < NF; ++q)
#pragma unroll
    for (int k = 0; k < KU; ++k) {
      const int t = threadIdx.x + k * kBlock;
      if (t < NU) su[q][t / UX][t % UX] = ureg[q][k];
    }
  __syncthreads();
  const R gsx = R(g.sx), gsy = R(g.sy);
  const R tx0 = R(g.tx0), tx1 = R(g.tx1), ty0 = R(g.ty0), ty1 = R(g.ty1);
#pragma unroll
  for (int q = 0; q < NF; ++q) {
    // a boundary face's value: the inlet's, or the cell's (zero derivative)
    const double* bv = fl.bval[q];
    auto bface = [&](int s, int pos, R uc) {
      return (IN && bv && fl.inlet[s]) ? R(bv[BoundaryIndex(g, s, pos)]) : uc;
    };
    // the column's values at tile rows lj0 - 2 .. lj0 + S + 1
    double col[S + 4];
#pragma unroll
    for (int r = 0; r < S + 4; ++r) col[r] = su[q][lj0 + r][li + 2];
    // y-gradients at rows j0 + lj0 - 1 .. j0 + lj0 + S (zero outside the grid)
    R gy[S + 2];
#pragma unroll
    for (int r = 0; r < S + 2; ++r) {
      const int jj = j0 + lj0 - 1 + r;
      R gv = R(0);
      if (i < nx && jj >= 0 && jj < ny) {
        const R uc = R(col[r + 1]);
        const R fs = jj > 0 ? MidR<R>(R(col[r]), uc) : bface(kBottom, i, uc);
        const R fn = jj < ny - 1 ? MidR<R>(uc, R(col[r + 2])) : bface(kTop, i, uc);
        R syv = (-gsy) * fs;
        syv += gsy * fn;
        gv = OverVR<R>(g, syv);
      }
      gy[r] = gv;
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int lj = lj0 + s, j = j0 + lj;
      if (i >= nx || j >= ny) continue;
      const int c = i + nx * j;
      // the row's values at i - 2 .. i + 2
      const R uW2 = R(su[q][lj + 2][li]), uW = R(su[q][lj + 2][li + 1]);
      const R uE = R(su[q][lj + 2][li + 3]), uE2 = R(su[q][lj + 2][li + 4]);
      const double uc64 = col[s + 2];
      const R uc = R(uc64);
      // x-gradient of cell ii of this row from its value u0 and its
      // neighbours' um, up (zero outside the grid)
      auto gx = [&](int ii, R um, R u0, R up) {
        R gv = R(0);
        if (ii >= 0 && ii < nx) {
          const R fw = ii > 0 ? MidR<R>(um, u0) : bface(kLeft, j, u0);
          const R fe = ii < nx - 1 ? MidR<R>(u0, up) : bface(kRight, j, u0);
          R sxv = (-gsx) * fw;
          sxv += gsx * fe;
          gv = OverVR<R>(g, sxv);
        }
        return gv;
      };
      const R gxc = gx(i, uW, uc, uE), gyc = gy[s + 1];
      const R vw = vw_[s], ve = ve_[s], vs = vy_[s], vn = vy_[s + 1];
      R uw = i > 0 ? uc : bface(kLeft, j, uc), ue = i < nx - 1 ? uc : bface(kRight, j, uc);
      R us_ = j > 0 ? uc : bface(kBottom, i, uc), uN = j < ny - 1 ? uc : bface(kTop, i, uc);
      if (i > 0) uw = SuperbeeFace<R>(uW, uc, gx(i - 1, uW2, uW, uc), gxc, tx0, tx1, vw);
      if (i < nx - 1) ue = SuperbeeFace<R>(uc, uE, gxc, gx(i + 1, uc, uE, uE2), tx0, tx1, ve);
      if (j > 0) us_ = SuperbeeFace<R>(R(col[s + 1]), uc, gy[s], gyc, ty0, ty1, vs);
      if (j < ny - 1) uN = SuperbeeFace<R>(uc, R(col[s + 3]), gyc, gy[s + 2], ty0, ty1, vn);
      double fsum = 0.;
      fsum += double(uw * vw * R(-1));
      fsum += double(ue * ve * R(1));
      fsum += double(us_ * vs * R(-1));
      fsum += double(uN * vn * R(1));
      fl.un[q][c] = uc64 + OverV(g, -dt) * fsum;
    }
  }
}

__global__ void __launch_bounds__(kBlock)
KProperties(Geo g, Phases ph, const double* __restrict__ pd0, const double* __restrict__ pd1,
            const double* __restrict__ pd2, double* __restrict__ a0, double* __restrict__ a1,
            double* __restrict__ a2, double* __restrict__ rho, double* __restrict__ mu,
            double* __restrict__ fx, double* __restrict__ fy) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nx * g.ny) return;
  double al[3] = {0, 0, 0};
  const double* pd[3] = {pd0, pd1, pd2};
  double sum = 0.;
  for (int i = 0; i < ph.n; ++i) {
    al[i] = pd[i][c] / ph.rho[i];
    sum += al[i];
  }
  for (int i = 0; i < ph.n; ++i) al[i] /= sum;
  double r = 0., m = 0.;
  for (int i = 0; i < ph.n; ++i) r += ph.rho[i] * al[i];
  for (int i = 0; i < ph.n; ++i) m += ph.mu[i] * al[i];
  a0[c] = al[0];
  if (ph.n > 1) a1[c] = al[1];
  if (ph.n > 2) a2[c] = al[2];
  rho[c] = r;
  mu[c] = m;
  fx[c] = ph.grav[0] * r + ph.force[0];
  fy[c] = ph.grav[1] * r + ph.force[1];
}

// One pass of GetSmoothField (solver.hpp:636-656): average of the
// zero-derivative face interpolation.
__global__ void __launch_bounds__(kBlock)
KSmooth(Geo g, const double* __restrict__ u, double* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nx * g.ny) return;
  double f[4];
  CellFaces<1>(g, u, c % g.nx, c / g.nx, nullptr, f);
  double s = 0.;
  s += f[0];
  s += f[1];
  s += f[2];
  s += f[3];
  out[c] = s / 4.;
}

// Two passes of GetSmoothField in one launch.  The block loads its 64 x 8
// tile with a two-cell halo into LDS, forms the first pass on the tile plus
// a one-cell halo, then the second pass on the tile.  Each value is formed
// with the operations of KSmooth, so the result is bit for bit that of two
// KSmooth launches.  `out` must not alias `u`.
constexpr int kSmTX = 64, kSmTY = 8;

__device__ __forceinline__ double SmoothAt(int nx, int ny, int i, int j, double uc, double uw,
                                           double ue, double us, double un) {
  const double fw = i > 0 ? Mid(uw, uc) : uc;
  const double fe = i < nx - 1 ? Mid(uc, ue) : uc;
  const double fs = j > 0 ? Mid(us, uc) : uc;
  const double fn = j < ny - 1 ? Mid(uc, un) : uc;
  double s = 0.;
  s += fw;
  s += fe;
  s += fs;
  s += fn;
  return s / 4.;
}

__global__ void __launch_bounds__(kBlock)
KSmooth2(Geo g, const double* __restrict__ u, double* __restrict__ out) {
  constexpr int UX = kSmTX + 4, UY = kSmTY + 4, PX = kSmTX + 2, PY = kSmTY + 2;
  __shared__ double su[UY][UX];
  __shared__ double sp[PY][PX];
  const int nx = g.nx, ny = g.ny;
  const int ntx = (nx + kSmTX - 1) / kSmTX;
  const int i0 = (blockIdx.x % ntx) * kSmTX, j0 = (blockIdx.x / ntx) * kSmTY;
  constexpr int NU = UX * UY, KU = (NU + kBlock - 1) / kBlock;
  double ureg[KU];
#pragma unroll
  for (int k = 0; k < KU; ++k) {
    const int t = threadIdx.x + k * kBlock;
    const int ii = i0 + t % UX - 2, jj = j0 + t / UX - 2;
    ureg[k] = (t < NU && ii >= 0 && ii < nx && jj >= 0 && jj < ny) ? u[ii + nx * jj] : 0.;
  }
#pragma unroll
  for (int k = 0; k < KU; ++k) {
    const int t = threadIdx.x + k * kBlock;
    if (t < NU) su[t / UX][t % UX] = ureg[k];
  }
  __syncthreads();
  // first pass on cells i0-1 .. i0+64, j0-1 .. j0+8 inside the grid
  for (int t = threadIdx.x; t < PX * PY; t += kBlock) {
    const int pj = t / PX, pi = t % PX;
    const int ii = i0 + pi - 1, jj = j0 + pj - 1;
    double v = 0.;
    if (ii >= 0 && ii < nx && jj >= 0 && jj < ny)
      v = SmoothAt(nx, ny, ii, jj, su[pj + 1][pi + 1], su[pj + 1][pi], su[pj + 1][pi + 2],
                   su[pj][pi + 1], su[pj + 2][pi + 1]);
    sp[pj][pi] = v;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kSmTX * kSmTY; t += kBlock) {
    const int lj = t / kSmTX, li = t % kSmTX;
    const int i = i0 + li, j = j0 + lj;
    if (i >= nx || j >= ny) continue;
    out[i + nx * j] = SmoothAt(nx, ny, i, j, sp[lj + 1][li + 1], sp[lj + 1][li],
                               sp[lj + 1][li + 2], sp[lj][li + 1], sp[lj + 2][li + 1]);
  }
}

// Initial partial density of phase 0 compensating the others (hydro2d.hpp:556-565).
__global__ void __launch_bounds__(kBlock)
KPhase0Init(Geo g, Phases ph, const double* __restrict__ pd1, const double* __restrict__ pd2,
            double* __restrict__ pd0) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nx * g.ny) return;
  double vs = 0.;
  if (ph.n > 1) vs += pd1[c] / ph.rho[1];
  if (ph.n > 2) vs += pd2[c] / ph.rho[2];
  pd0[c] = (1. - vs) * ph.rho[0];
}

// NaN test of FinishStep (fluid.hpp:1164-1166, conv_diff.hpp:255-257):
// IsNan(v) is !(v * 0 == 0) (solver.hpp:18-20).  Bit 1 of the flag: the
// first field (pressure); bit 2: one of the other two (velocity).
__global__ void __launch_bounds__(kBlock)
KNanCheck(const double* __restrict__ a, const double* __restrict__ b,
          const double* __restrict__ c3, int n, int* __restrict__ flag) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const bool na = !(a[c] * 0. == 0.);
  const bool nbc = !(b[c] * 0. + c3[c] * 0. == 0.);
  if (na || nbc) atomicOr(flag, (na ? 1 : 0) | (nbc ? 2 : 0));
}

// CalcStat partial sums (hydro2d.hpp:1432-1467) for one phase: per block,
// volume, centroid, velocity sums and pd min/max; a second pass folds the
// block partials in a fixed order.
__global__ void __launch_bounds__(kBlock)
KStatsPartial(Geo g, const double* __restrict__ xn, const double* __restrict__ yn,
              const double* __restrict__ alpha, const double* __restrict__ pd,
              const double* __restrict__ u0, const double* __restrict__ u1,
              double* __restrict__ part) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double v[7] = {0, 0, 0, 0, 0, 1e10, -1e10};
  if (c < g.nx * g.ny) {
    const int i = c % g.nx, j = c / g.nx;
    double cx = 0., cy = 0.;
    cx += xn[i]; cy += yn[j];
    cx += xn[i + 1]; cy += yn[j];
    cx += xn[i + 1]; cy += yn[j + 1];
    cx += xn[i]; cy += yn[j + 1];
    cx /= 4.;
    cy /= 4.;
    const double a = alpha[c];
    const double w = a * g.V;
    v[0] = a * g.V;
    v[1] = cx * w;
    v[2] = cy * w;
    v[3] = u0[c] * w;
    v[4] = u1[c] * w;
    v[5] = pd[c];
    v[6] = pd[c];
  }
  __shared__ double red[7][kBlock / 64];
  for (int q = 0; q < 7; ++q) {
    double x = q < 5 ? WaveSum(v[q]) : (q == 5 ? WaveMin(v[q]) : WaveMax(v[q]));
    if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x < 7) {
    const int q = threadIdx.x;
    double x = red[q][0];
    for (int w = 1; w < kBlock / 64; ++w)
      x = q < 5 ? x + red[q][w] : (q == 5 ? fmin(x, red[q][w]) : fmax(x, red[q][w]));
    part[blockIdx.x * 7 + q] = x;
  }
}

__global__ void KStatsFinal(const double* __restrict__ part, int nblocks, double* __restrict__ out) {
  const int q = threadIdx.x;
  if (q >= 7) return;
  double x = q < 5 ? 0. : (q == 5 ? 1e10 : -1e10);
  for (int b = 0; b < nblocks; ++b) {
    double y = part[b * 7 + q];
    x = q < 5 ? x + y : (q == 5 ? fmin(x, y) : fmax(x, y));
  }
  out[q] = x;
}

// ---- SIMPLER (fluid.hpp:1059-1151) -------------------------------------------
// The momentum rows of the iteration evaluated on the velocity change
// u_curr - u_prev (Expression::Evaluate, terms S, W, C, E, N), plus the
// restored force minus the pressure gradient, per component.
__global__ void __launch_bounds__(kBlock)
KSimplerCells(Geo g, Rows a, const double* __restrict__ k0, const double* __restrict__ k1,
              const double* __restrict__ uc0, const double* __restrict__ uc1,
              const double* __restrict__ up0, const double* __restrict__ up1,
              const double* __restrict__ frx, const double* __restrict__ fry,
              const double* __restrict__ gpx, const double* __restrict__ gpy,
              double* __restrict__ ev0, double* __restrict__ ev1) {
  const int nx = g.nx, ny = g.ny;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  const int i = c % nx, j = c / nx;
  const Cross c0 = LoadCross(uc0, nx, ny, i, j), p0 = LoadCross(up0, nx, ny, i, j);
  const Cross c1 = LoadCross(uc1, nx, ny, i, j), p1 = LoadCross(up1, nx, ny, i, j);
  const double as = a.s[c], aw = a.w[c], ac = a.c[c], ae = a.e[c], an = a.n[c];
  const double kk0 = k0[c], kk1 = k1[c];
  const double f0 = frx[c], f1 = fry[c], g0 = gpx[c], g1 = gpy[c];
  double r0 = kk0, r1 = kk1;
  if (j > 0) {
    r0 += (c0.s - p0.s) * as;
    r1 += (c1.s - p1.s) * as;
  }
  if (i > 0) {
    r0 += (c0.w - p0.w) * aw;
    r1 += (c1.w - p1.w) * aw;
  }
  r0 += (c0.c - p0.c) * ac;
  r1 += (c1.c - p1.c) * ac;
  if (i < nx - 1) {
    r0 += (c0.e - p0.e) * ae;
    r1 += (c1.e - p1.e) * ae;
  }
  if (j < ny - 1) {
    r0 += (c0.n - p0.n) * an;
    r1 += (c1.n - p1.n) * an;
  }
  ev0[c] = r0 + (f0 - g0);
  ev1[c] = r1 + (f1 - g1);
}

// The second pressure right-hand side: per inner face
//   ((ev_f - f_f) . S) / a_P,f + (flux - u_f . S) / rhie_chow_factor
// (face means of the cell values), the outward sum over the cell's faces,
// negated, then the pressure row evaluated at the current pressure; a fixed
// cell keeps -value and its neighbours' rows stay without its term.
__global__ void __launch_bounds__(kBlock)
KSimplerRhs(Geo g, const double* __restrict__ ev0, const double* __restrict__ ev1,
            const double* __restrict__ ffex, const double* __restrict__ ffey,
            const double* __restrict__ ap, const double* __restrict__ flux,
            const double* __restrict__ u0, const double* __restrict__ u1, double rc, Rows pr,
            const double* __restrict__ p, int pinned, double pval, double* __restrict__ k) {
  const int nx = g.nx, ny = g.ny;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  const int i = c % nx, j = c / nx;
  const bool inner[4] = {i > 0, i < nx - 1, j > 0, j < ny - 1};
  const int face[4] = {XFace(g, i, j), XFace(g, i + 1, j), YFace(g, i, j), YFace(g, i, j + 1)};
  const int nb[4] = {c - 1, c + 1, c - nx, c + nx};
  const double outf[4] = {-1., 1., -1., 1.};
  const Cross e0 = LoadCross(ev0, nx, ny, i, j), e1 = LoadCross(ev1, nx, ny, i, j);
  const Cross a = LoadCross(ap, nx, ny, i, j);
  const Cross v0 = LoadCross(u0, nx, ny, i, j), v1 = LoadCross(u1, nx, ny, i, j);
  const Cross pp = LoadCross(p, nx, ny, i, j);
  double fl[4], fe[4];
  for (int q = 0; q < 4; ++q) {
    fl[q] = flux[face[q]];
    fe[q] = q < 2 ? ffex[face[q]] : ffey[face[q]];
  }
  const double enb[4] = {e0.w, e0.e, e1.s, e1.n}, ec[4] = {e0.c, e0.c, e1.c, e1.c};
  const double anb[4] = {a.w, a.e, a.s, a.n};
  const double unb[4] = {v0.w, v0.e, v1.s, v1.n}, uc[4] = {v0.c, v0.c, v1.c, v1.c};
  double sum = 0.;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double r = 0.;
    if (inner[q]) {
      const bool self_m = (q == 1 || q == 3);  // the cell is the face's lower-index side
      const double S = q < 2 ? g.sx : g.sy;
      const double em = self_m ? ec[q] : enb[q], ep = self_m ? enb[q] : ec[q];
      const double am = self_m ? a.c : anb[q], apl = self_m ? anb[q] : a.c;
      const double um = self_m ? uc[q] : unb[q], up = self_m ? unb[q] : uc[q];
      const double evf = Mid(em, ep) - fe[q];
      r = (evf * S) / Mid(am, apl) + (fl[q] - Mid(um, up) * S) / rc;
    }
    sum += r * outf[q];
  }
  double kk = -sum;
  if (c == pinned) {
    kk = -pval;
    kk += pp.c * pr.c[c];
  } else {
    if (inner[2] && nb[2] != pinned) kk += pp.s * pr.s[c];
    if (inner[0] && nb[0] != pinned) kk += pp.w * pr.w[c];
    kk += pp.c * pr.c[c];
    if (inner[1] && nb[1] != pinned) kk += pp.e * pr.e[c];
    if (inner[3] && nb[3] != pinned) kk += pp.n * pr.n[c];
  }
  k[c] = kk;
}

// Per-phase cell fields and a per-phase constant (density, conductivity).
struct PhaseCells {
  const double* alpha[3];
  double density[3];
};

// ---- inlets and outlets (fluid.hpp:542-600, hydro2d.hpp:565-583) ------------
struct SideKinds {
  int kind[4];  // hg_bc_kind per side
};

// The geometry of boundary face b: side, position along it, face, cell.
__device__ __forceinline__ void BoundaryFace(const Geo& g, int b, int& s, int& q, int& face,
                                             int& cell) {
  if (b < 2 * g.ny) {
    s = b < g.ny ? kLeft : kRight;
    q = b < g.ny ? b : b - g.ny;
    face = XFace(g, s == kLeft ? 0 : g.nx, q);
    cell = (s == kLeft ? 0 : g.nx - 1) + g.nx * q;
  } else {
    const int r = b - 2 * g.ny;
    s = r < g.nx ? kBottom : kTop;
    q = r < g.nx ? r : r - g.nx;
    face = YFace(g, q, s == kBottom ? 0 : g.ny);
    cell = q + g.nx * (s == kBottom ? 0 : g.ny - 1);
  }
}

// u . S of a boundary face (x-faces S = (sx, -0), y-faces S = (+0, sy)), in
// Vect::dot's order.
__device__ __forceinline__ double BoundaryDot(const Geo& g, int s, double ux, double uy) {
  const bool xf = s < 2;
  const double Sx = xf ? g.sx : 0., Sy = xf ? -0. : g.sy;
  double d = 0.;
  d += ux * Sx;
  d += uy * Sy;
  return d;
}

// Inlet values of the phases: the adjacent cells' partial densities.
__global__ void KGatherInlet(Geo g, SideKinds sk, PhaseCells pd, int np, double* __restrict__ pdb) {
  const int nb = 2 * (g.nx + g.ny);
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  int s, q, face, cell;
  BoundaryFace(g, b, s, q, face, cell);
  for (int p = 0; p < np; ++p) pdb[p * nb + b] = sk.kind[s] == 1 ? pd.alpha[p][cell] : 0.;
}

// Outlet faces take the adjacent cell's velocity; per face, the terms of
// the outflow (u . S toward the outside) and area, or of the inflow.
__global__ void KOutletTerms(Geo g, SideKinds sk, const double* __restrict__ u0,
                             const double* __restrict__ u1, double* __restrict__ bv0,
                             double* __restrict__ bv1, double* __restrict__ term,
                             double* __restrict__ area) {
  const int nb = 2 * (g.nx + g.ny);
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  int s, q, face, cell;
  BoundaryFace(g, b, s, q, face, cell);
  (void)face;
  double t = 0., a = 0.;
  if (sk.kind[s] == 2) {
    const double ux = u0[cell], uy = u1[cell];
    bv0[b] = ux;
    bv1[b] = uy;
    t = BoundaryDot(g, s, ux, uy) * SideFactor(s);
    a = s < 2 ? g.ax : g.ay;
  } else if (sk.kind[s] == 1) {
    t = BoundaryDot(g, s, bv0[b], bv1[b]) * (-SideFactor(s));
  }
  term[b] = t;
  area[b] = a;
}

// The sums in the reference's face order (rows of left / right x-faces,
// then the bottom and top y-faces), one thread, staged through LDS; then
// corr = (inflow - outflow) / outlet area.
__global__ void __launch_bounds__(kBlock)
KOutletSum(Geo g, SideKinds sk, const double* __restrict__ term,
           const double* __restrict__ area, double* __restrict__ corr) {
  const int nb = 2 * (g.nx + g.ny);
  __shared__ double st[kBlock], sa[kBlock];
  __shared__ int sk_[kBlock];
  double in = 0., out = 0., ar = 0.;
  for (int k0 = 0; k0 < nb; k0 += kBlock) {
    const int k = k0 + threadIdx.x;
    if (k < nb) {
      const int b = k < 2 * g.ny ? (k & 1) * g.ny + (k >> 1) : k;
      int s, q, face, cell;
      BoundaryFace(g, b, s, q, face, cell);
      st[threadIdx.x] = term[b];
      sa[threadIdx.x] = area[b];
      sk_[threadIdx.x] = sk.kind[s];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int n = nb - k0 < kBlock ? nb - k0 : kBlock;
      for (int t = 0; t < n; ++t) {
        if (sk_[t] == 2) {
          out += st[t];
          ar += sa[t];
        } else if (sk_[t] == 1) {
          in += st[t];
        }
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    in += 0.;  // the volume sources (zero on this path) times the cell volumes
    corr[0] = (in - out) / ar;
  }
}

// Outlet velocity += normal * corr * factor (fluid.hpp:583-599).
__global__ void KOutletApply(Geo g, SideKinds sk, const double* __restrict__ corr,
                             double* __restrict__ bv0, double* __restrict__ bv1) {
  const int nb = 2 * (g.nx + g.ny);
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  int s, q, face, cell;
  BoundaryFace(g, b, s, q, face, cell);
  if (sk.kind[s] != 2) return;
  const bool xf = s < 2;
  const double A = xf ? g.ax : g.ay;
  const double nx_ = (xf ? g.sx : 0.) / A, ny_ = (xf ? -0. : g.sy) / A;
  const double c = corr[0], factor = SideFactor(s);
  bv0[b] = bv0[b] + (nx_ * c) * factor;
  bv1[b] = bv1[b] + (ny_ * c) * factor;
}

// ---- heat: the temperature's implicit convection-diffusion iteration ---------
// (HeatSolver::MakeIteration, heat.hpp:66-75, on ConvectionDiffusionScalar-
// Implicit::MakeIteration, conv_diff.hpp:130-243) with the scaling 1, zero
// source, the step's volume flux and the face conductivity interpolated with
// zero normal derivative on the boundary.  Boundary face b (left j, right
// ny + j, bottom 2 ny + i, top 2 ny + nx + i) holds either a fixed value
// (kind 0) or a fixed normal derivative (kind 1).

// Cell conductivity (GetVolumeAveraged of conductivity_i, hydro2d.hpp:1236-1246)
// and the gradient of the temperature's face values (Interpolate with the
// heat conditions, solver.hpp:392-470, then Gradient, solver.hpp:658-677).
__global__ void __launch_bounds__(kBlock)
KHeatGrad(Geo g, HeatFaces hb, PhaseCells pc, int np, const double* __restrict__ t,
          double* __restrict__ cond, double* __restrict__ gx, double* __restrict__ gy) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nx * g.ny) return;
  const int i = c % g.nx, j = c / g.nx;
  const Cross x = LoadCross(t, g.nx, g.ny, i, j);
  double a[3] = {0., 0., 0.};
  for (int p = 0; p < 3; ++p)
    if (p < np) a[p] = pc.alpha[p][c];
  double k = 0.;
  for (int p = 0; p < 3; ++p)
    if (p < np) k += pc.density[p] * a[p];
  cond[c] = k;
  double f[4];
  CellFacesV<1>(g, x, i, j, nullptr, f);
  const int sides[4] = {kLeft, kRight, kBottom, kTop};
  const bool bnd[4] = {i == 0, i == g.nx - 1, j == 0, j == g.ny - 1};
  for (int q = 0; q < 4; ++q) {
    if (!bnd[q]) continue;
    const int s = sides[q];
    const int b = BoundaryIndex(g, s, s < 2 ? j : i);
    const double v = hb.val[b];
    f[q] = hb.kind[b] == 0 ? v : x.c + v * (g.ex_dist[s] * SideFactor(s));
  }
  GradFromFaces(g, f, gx[c], gy[c]);
}

__global__ void __launch_bounds__(kBlock)
KHeatRows(Geo g, HeatFaces hb, HeatIn in, double* __restrict__ rs_, double* __restrict__ rw_,
          double* __restrict__ rc_, double* __restrict__ re_, double* __restrict__ rn_,
          double* __restrict__ rk_) {
  const int nx = g.nx, ny = g.ny;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  const int i = c % nx, j = c / nx;
  const bool hw = i > 0, he = i < nx - 1, hs = j > 0, hn = j < ny - 1;
  const int fw = XFace(g, i, j), fe = XFace(g, i + 1, j);
  const int fs = YFace(g, i, j), fn = YFace(g, i, j + 1);
  const double V = g.V;
  // inputs first (neighbours at clamped indices), selected below
  const Cross t = LoadCross(in.t, nx, ny, i, j);
  const Cross kc = LoadCross(in.cond, nx, ny, i, j);
  const Cross gxc = LoadCross(in.gx, nx, ny, i, j);
  const Cross gyc = LoadCross(in.gy, nx, ny, i, j);
  const double Ff[4] = {in.flux[fw], in.flux[fe], in.flux[fs], in.flux[fn]};
  const double ttp = in.ttp[c], ttc = in.ttc[c];
  const bool inner[4] = {hw, he, hs, hn};
  const double knb[4] = {kc.w, kc.e, kc.s, kc.n};
  // the face-normal gradient component at the neighbour and at the cell
  const double gnb[4] = {gxc.w, gxc.e, gyc.s, gyc.n};
  const double gcc[4] = {gxc.c, gxc.c, gyc.c, gyc.c};
  const double outf[4] = {-1., 1., -1., 1.};
  const int side[4] = {kLeft, kRight, kBottom, kTop};
  double Cc = 0., Dc = 0., Ck = 0., Dk = 0., Cnb[4] = {0, 0, 0, 0}, Dnb[4] = {0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const bool xface = q < 2;
    const double A = xface ? g.ax : g.ay;
    const double F = Ff[q];
    const double o = outf[q];
    const bool self_m = (q == 1 || q == 3);
    if (inner[q]) {
      // face conductivity: the mean of the two cells (lower index first)
      const double D = Mid(self_m ? kc.c : knb[q], self_m ? knb[q] : kc.c);
      const double gm = self_m ? gcc[q] : gnb[q], gp = self_m ? gnb[q] : gcc[q];
      double am, ap, ak;
      if (F > 1e-10) {
        am = 1.;
        ap = 0.;
        ak = -(xface ? gm * g.tx0 : gm * g.ty0);
      } else if (F < -1e-10) {
        am = 0.;
        ap = 1.;
        ak = -(xface ? gp * g.tx1 : gp * g.ty1);
      } else {
        am = 0.5;
        ap = 0.5;
        ak = 0.;
      }
      const double cam = am * F, cap = ap * F;
      const double al = 1. / (xface ? g.dnx : g.dny);
      const double dcm = ((-al) * (-D)) * A;
      const double dcp = (al * (-D)) * A;
      const double dck = (0. * (-D)) * A;
      Cc += (self_m ? cam : cap) * o;
      Cnb[q] = (self_m ? cap : cam) * o;
      Dc += (self_m ? dcm : dcp) * o;
      Dnb[q] = (self_m ? dcp : dcm) * o;
      Ck += (ak * F) * o;
      Dk += dck * o;
    } else {
      const int s = side[q];
      const int b = BoundaryIndex(g, s, s < 2 ? j : i);
      const double v = hb.val[b];
      // the boundary face's conductivity: the cell's, zero normal derivative
      const double D = kc.c + 0. * (g.ex_dist[s] * SideFactor(s));
      if (hb.kind[b] == 0) {
        // value fixed: convective constant v F; diffusive -al on the cell
        const double al = g.bal[s];
        Ck += (v * F) * o;
        Dc += (((-al) * (-D)) * A) * o;
        Dk += (((al * v) * (-D)) * A) * o;
      } else {
        // derivative fixed: the value is the cell's + v * signed distance
        Cc += (1. * F) * o;
        Ck += (((g.ex_dist[s] * SideFactor(s)) * v) * F) * o;
        Dk += ((v * (-D)) * A) * o;
      }
    }
  }
  const double cen = OverV(g, Dc) + ((OverV(g, Cc) + in.bdf2) * 1.);
  double nb[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) nb[q] = OverV(g, Dnb[q]) + (OverV(g, Cnb[q]) * 1.);
  const double un = in.bdf0 * ttp + in.bdf1 * ttc;
  double k = (OverV(g, Dk) + ((OverV(g, Ck) + un) * 1.)) + (-0.);
  if (hs) k += t.s * nb[2];
  if (hw) k += t.w * nb[0];
  k += t.c * cen;
  if (he) k += t.e * nb[1];
  if (hn) k += t.n * nb[3];
  rs_[c] = hs ? nb[2] : 0.;
  rw_[c] = hw ? nb[0] : 0.;
  rc_[c] = cen / in.relax;
  re_[c] = he ? nb[1] : 0.;
  rn_[c] = hn ? nb[3] : 0.;
  rk_[c] = k;
}

// Boundary in/out integrals of CalcStat (hydro2d.hpp:1474-1508): per
// phase, the volume flux into the domain through boundary faces times the
// cell's volume fraction (and out of it), then the four running totals +=
// flux * dt (volume and mass in and out).  The boundary faces are taken
// left, right (x-faces i = 0, nx), bottom, top (y-faces j = 0, ny); the flux
// toward the inner cell is +flux on left and bottom faces, -flux on right
// and top faces (GetValidNeighbourCellId 1 / 0).  One thread per boundary
// face; per-workgroup partial sums, folded by one workgroup.
__global__ void __launch_bounds__(kBlock)
KBoundaryFluxPartial(Geo g, const double* __restrict__ flux, PhaseCells pc, int np,
                     double* __restrict__ part) {
  const int nb = 2 * (g.nx + g.ny);
  const int b = blockIdx.x * kBlock + threadIdx.x;
  double in[3] = {0., 0., 0.}, out[3] = {0., 0., 0.};
  if (b < nb) {
    int face, cell;
    double factor;
    if (b < 2 * g.ny) {
      const bool left = b < g.ny;
      const int j = left ? b : b - g.ny;
      face = XFace(g, left ? 0 : g.nx, j);
      cell = (left ? 0 : g.nx - 1) + g.nx * j;
      factor = left ? 1. : -1.;
    } else {
      const int r = b - 2 * g.ny;
      const bool bottom = r < g.nx;
      const int i = bottom ? r : r - g.nx;
      face = YFace(g, i, bottom ? 0 : g.ny);
      cell = i + g.nx * (bottom ? 0 : g.ny - 1);
      factor = bottom ? 1. : -1.;
    }
    double vf[3] = {0., 0., 0.};
    for (int p = 0; p < 3; ++p)
      if (p < np) vf[p] = pc.alpha[p][cell];
    const double to_cell = flux[face] * factor;
    for (int p = 0; p < 3; ++p) {
      if (to_cell >= 0) in[p] += to_cell * vf[p];
      else out[p] -= to_cell * vf[p];
    }
  }
  __shared__ double red[6][kBlock / 64];
  for (int q = 0; q < 6; ++q) {
    const double x = WaveSum(q < 3 ? in[q] : out[q - 3]);
    if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double x = red[threadIdx.x][0];
    for (int w = 1; w < kBlock / 64; ++w) x += red[threadIdx.x][w];
    part[blockIdx.x * 6 + threadIdx.x] = x;
  }
}

__global__ void KBoundaryFluxFinal(const double* __restrict__ part, int nblocks, PhaseCells pc,
                                   int np, double dt, double* __restrict__ acc) {
  const int p = threadIdx.x;
  if (p >= np) return;
  double vin = 0., vout = 0.;
  for (int k = 0; k < nblocks; ++k) {
    vin += part[k * 6 + p];
    vout += part[k * 6 + 3 + p];
  }
  const double min = vin * pc.density[p], mout = vout * pc.density[p];
  acc[4 * p + 0] += vin * dt;
  acc[4 * p + 1] += vout * dt;
  acc[4 * p + 2] += min * dt;
  acc[4 * p + 3] += mout * dt;
}

// FluidSimple::GetAutoTimeStep (fluid.hpp:1191-1206): the smallest
// |V / flux| over every cell's four faces with a nonzero time_curr flux,
// 1e10 if there is none.  A minimum is exact in any order.
__global__ void __launch_bounds__(kBlock)
KAutoDtPartial(Geo g, const double* __restrict__ flux, double* __restrict__ part) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double dt = 1e10;
  if (c < g.nx * g.ny) {
    const int i = c % g.nx, j = c / g.nx;
    const double f[4] = {flux[XFace(g, i, j)], flux[XFace(g, i + 1, j)], flux[YFace(g, i, j)],
                         flux[YFace(g, i, j + 1)]};
    for (int k = 0; k < 4; ++k)
      if (f[k] != 0.) {
        const double x = fabs(g.V / f[k]);
        dt = x < dt ? x : dt;
      }
  }
  __shared__ double red[kBlock / 64];
  const double w = WaveMin(dt);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = red[0];
    for (int q = 1; q < kBlock / 64; ++q) x = fmin(x, red[q]);
    part[blockIdx.x] = x;
  }
}

__global__ void __launch_bounds__(kBlock)
KAutoDtFinal(const double* __restrict__ part, int nblocks, double* __restrict__ out) {
  double x = 1e10;
  for (int b = threadIdx.x; b < nblocks; b += kBlock) x = fmin(x, part[b]);
  __shared__ double red[kBlock / 64];
  const double w = WaveMin(x);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    double y = red[0];
    for (int q = 1; q < kBlock / 64; ++q) y = fmin(y, red[q]);
    out[0] = y;
  }
}

}  // namespace

// ---- launchers --------------------------------------------------------------

template <class T>
void ForceFacesT(hipStream_t st, const Geo& g, const double* fx, const double* fy,
                 const double* mu, T* ffex, T* ffey, T* ffmu) {
  long nf = long(g.nfx) + long(g.nx) * (g.ny + 1);
  KForceFaces<T><<<Grid(nf), kBlock, 0, st>>>(g, fx, fy, mu, ffex, ffey, ffmu);
}
void LaunchForceFaces(hipStream_t st, const Geo& g, const double* fx, const double* fy,
                      const double* mu, double* ffex, double* ffey, double* ffmu) {
  ForceFacesT<double>(st, g, fx, fy, mu, ffex, ffey, ffmu);
}
void LaunchForceFaces(hipStream_t st, const Geo& g, const double* fx, const double* fy,
                      const double* mu, float* ffex, float* ffey, float* ffmu) {
  ForceFacesT<float>(st, g, fx, fy, mu, ffex, ffey, ffmu);
}
void LaunchRestoredForce(hipStream_t st, const Geo& g, const double* ffex, const double* ffey,
                         double* frx, double* fry) {
  KRestoredForce<double><<<Grid(long(g.nx) * g.ny), kBlock, 0, st>>>(g, ffex, ffey, frx, fry);
}
void LaunchRestoredForce(hipStream_t st, const Geo& g, const float* ffex, const float* ffey,
                         float* frx, float* fry) {
  KRestoredForce<float><<<Grid(long(g.nx) * g.ny), kBlock, 0, st>>>(g, ffex, ffey, frx, fry);
}
void LaunchFacesDeriv0Pair(hipStream_t st, const Geo& g, const double* ax, const double* ay,
                           double* fx, double* fy) {
  long nf = long(g.nfx) + long(g.nx) * (g.ny + 1);
  KFacesDeriv0Pair<double><<<Grid(nf), kBlock, 0, st>>>(g, ax, ay, fx, fy);
}
void LaunchFacesDeriv0Pair(hipStream_t st, const Geo& g, const float* ax, const float* ay,
                           float* fx, float* fy) {
  long nf = long(g.nfx) + long(g.nx) * (g.ny + 1);
  KFacesDeriv0Pair<float><<<Grid(nf), kBlock, 0, st>>>(g, ax, ay, fx, fy);
}
template <class T>
void CellGradsT(hipStream_t st, const Geo& g, const Bc& bc, const double* p, const double* u0,
                const double* u1, T* gpx, T* gpy, T* g0x, T* g0y, T* g1x, T* g1y) {
  if (bc.fvel[0] || bc.fvel[1])
    KCellGrads<true, T><<<Grid(long(g.nx) * g.ny), kBlock, 0, st>>>(g, bc, p, u0, u1, gpx, gpy,
                                                                   g0x, g0y, g1x, g1y);
  else
    KCellGrads<false, T><<<Grid(long(g.nx) * g.ny), kBlock, 0, st>>>(g, bc, p, u0, u1, gpx, gpy,
                                                                    g0x, g0y, g1x, g1y);
}
void LaunchCellGrads(hipStream_t st, const Geo& g, const Bc& bc, const double* p,
                     const double* u0, const double* u1, double* gpx, double* gpy, double* g0x,
                     double* g0y, double* g1x, double* g1y) {
  CellGradsT<double>(st, g, bc, p, u0, u1, gpx, gpy, g0x, g0y, g1x, g1y);
}
void LaunchCellGrads(hipStream_t st, const Geo& g, const Bc& bc, const double* p,
                     const double* u0, const double* u1, float* gpx, float* gpy, float* g0x,
                     float* g0y, float* g1x, float* g1y) {
  CellGradsT<float>(st, g, bc, p, u0, u1, gpx, gpy, g0x, g0y, g1x, g1y);
}
template <class T>
void MomentumT(hipStream_t st, const Geo& g, const Bc& bc, const MomentumArgsT<T>& a) {
  MomIn<T> in{a.u0,   a.u1,   a.g0x,  a.g0y,  a.g1x,  a.g1y,  a.gpx,  a.gpy,  a.frx, a.fry,
              a.ffmu, a.flux, a.rho,  a.utp0, a.utp1, a.utc0, a.utc1, a.bdf0, a.bdf1, a.bdf2,
              a.relax};
  MomOut<T> out{a.s, a.w, a.c, a.e, a.n, a.k0, a.k1, a.ap};
  KMomentum<T><<<Grid(long(g.nx) * g.ny), kBlock, 0, st>>>(g, bc, in, out);
}
void LaunchMomentum(hipStream_t st, const Geo& g, const Bc& bc, const MomentumArgs& a) {
  MomentumT<double>(st, g, bc, a);
}
void LaunchMomentum(hipStream_t st, const Geo& g, const Bc& bc, const MomentumArgsF& a) {
  MomentumT<float>(st, g, bc, a);
}
template <class T>
void RhieChowT(hipStream_t st, const Geo& g, const Bc& bc, const double* u0, const double* u1,
               const T* gpx, const T* gpy, const T* ffrx, const T* ffry, const T* ffex,
               const T* ffey, const double* p, const double* ap, double rc, double mvx,
               double mvy, double* fstar, double* fcoef) {
  long nf = long(g.nfx) + long(g.nx) * (g.ny + 1);
  KRhieChow<T><<<Grid(nf), kBlock, 0, st>>>(g, bc, u0, u1, gpx, gpy, ffrx, ffry, ffex, ffey, p,
                                            ap, rc, mvx, mvy, fstar, fcoef);
}
void LaunchRhieChow(hipStream_t st, const Geo& g, const Bc& bc, const double* u0,
                    const double* u1, const double* gpx, const double* gpy, const double* ffrx,
                    const double* ffry, const double* ffex, const double* ffey, const double* p,
                    const double* ap, double rc, double mvx, double mvy, double* fstar,
                    double* fcoef) {
  RhieChowT<double>(st, g, bc, u0, u1, gpx, gpy, ffrx, ffry, ffex, ffey, p, ap, rc, mvx, mvy,
                    fstar, fcoef);
}
void LaunchRhieChow(hipStream_t st, const Geo& g, const Bc& bc, const double* u0,
                    const double* u1, const float* gpx, const float* gpy, const float* ffrx,
                    const float* ffry, const float* ffex, const float* ffey, const double* p,
                    const double* ap, double rc, double mvx, double mvy, double* fstar,
                    double* fcoef) {
  RhieChowT<float>(st, g, bc, u0, u1, gpx, gpy, ffrx, ffry, ffex, ffey, p, ap, rc, mvx, mvy,
                   fstar, fcoef);
}
void LaunchPressureRhs(hipStream_t st, const Geo& g, const double* fstar, double* k,
                       bool negate) {
  KPressureRhs<<<Grid(long(g.nx) * g.ny), kBlock, 0, st>>>(g, fstar, k, negate ? 1 : 0);
}
void LaunchPressureRows(hipStream_t st, const Geo& g, const double* fstar, const double* fcoef,
                        int pinned, double pval, double* s, double* w, double* c, double* e,
                        double* n, double* k) {
  KPressureRows<<<Grid(long(g.nx) * g.ny), kBlock, 0, st>>>(g, fstar, fcoef, pinned, pval, s, w,
                                                             c, e, n, k);
}
void LaunchCorrectCells(hipStream_t st, const Geo& g, const double* pc, const double* ap,
                        const double* pprev, double alpha_p, double* pcurr, double* u0,
                        double* u1, const double* up0, const double* up1,
                        double* part, double* rs) {
  const long n = long(g.nx) * g.ny;
  KCorrectCells<<<Grid(n), kBlock, 0, st>>>(g, pc, ap, pprev, alpha_p, pcurr, u0, u1, up0, up1,
                                            part);
  KMaxFold<<<1, kFoldThreads, 0, st>>>(part, int((n + kBlock - 1) / kBlock), rs);
}
void LaunchCorrectFaces(hipStream_t st, const Geo& g, const double* fstar, const double* fcoef,
                        const double* pc, double* flux, float* flux32) {
  long nf = long(g.nfx) + long(g.nx) * (g.ny + 1);
  KCorrectFaces<<<Grid(nf), kBlock, 0, st>>>(g, fstar, fcoef, pc, flux, flux32);
}
void LaunchInitialFlux(hipStream_t st, const Geo& g, const Bc& bc, const double* u0,
                       const double* u1, double mvx, double mvy, double* flux) {
  long nf = long(g.nfx) + long(g.nx) * (g.ny + 1);
  KInitialFlux<<<Grid(nf), kBlock, 0, st>>>(g, bc, u0, u1, mvx, mvy, flux);
}
__global__ void __launch_bounds__(kBlock)
KToFloat(long n, const double* __restrict__ a, float* __restrict__ b) {
  const long i = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (i < n) b[i] = float(a[i]);
}
void LaunchToFloat(hipStream_t st, long n, const double* a, float* b) {
  KToFloat<<<Grid(n), kBlock, 0, st>>>(n, a, b);
}
template <class R>
void AdvStepT(hipStream_t st, const Geo& g, int nf, const double* const u[3], const R* flux,
              double dt, double* const un[3], const double* const bval[3], const int* inlet) {
  AdvFields fl{};
  for (int q = 0; q < nf; ++q) {
    fl.u[q] = u[q];
    fl.un[q] = un[q];
    fl.bval[q] = bval ? bval[q] : nullptr;
  }
  for (int s = 0; s < 4; ++s) fl.inlet[s] = inlet ? inlet[s] : 0;
  // IN: some field has inlet values on some inlet side (else every
  // boundary face takes the cell's value)
  bool any_in = false;
  for (int q = 0; q < nf; ++q)
    for (int s = 0; s < 4; ++s) any_in = any_in || (fl.bval[q] && fl.inlet[s]);
  const unsigned grid = unsigned(((g.nx + kAdvTX - 1) / kAdvTX) * ((g.ny + kAdvTY - 1) / kAdvTY));
  if (any_in) {
    if (nf == 1)
      KAdvStep<1, true, R><<<grid, kBlock, 0, st>>>(g, fl, flux, dt);
    else if (nf == 2)
      KAdvStep<2, true, R><<<grid, kBlock, 0, st>>>(g, fl, flux, dt);
    else
      KAdvStep<3, true, R><<<grid, kBlock, 0, st>>>(g, fl, flux, dt);
  } else {
    if (nf == 1)
      KAdvStep<1, false, R><<<grid, kBlock, 0, st>>>(g, fl, flux, dt);
    else if (nf == 2)
      KAdvStep<2, false, R><<<grid, kBlock, 0, st>>>(g, fl, flux, dt);
    else
      KAdvStep<3, false, R><<<grid, kBlock, 0, st>>>(g, fl, flux, dt);
  }
}
void LaunchAdvStep(hipStream_t st, const Geo& g, int nf, const double* const u[3],
                   const double* flux, double dt, double* const un[3],
                   const double* const bval[3], const int* inlet) {
  AdvStepT<double>(st, g, nf, u, flux, dt, un, bval, inlet);
}
void LaunchAdvStep(hipStream_t st, const Geo& g, int nf, const double* const u[3],
                   const float* flux, double dt, double* const un[3],
                   const double* const bval[3], const int* inlet) {
  AdvStepT<float>(st, g, nf, u, flux, dt, un, bval, inlet);
}
void LaunchProperties(hipStream_t st, const Geo& g, const Phases& ph, const double* const pd[3],
                      double* const alpha[3], double* rho, double* mu, double* fx, double* fy) {
  KProperties<<<Grid(long(g.nx) * g.ny), kBlock, 0, st>>>(g, ph, pd[0], pd[1], pd[2], alpha[0],
                                                           alpha[1], alpha[2], rho, mu, fx, fy);
}
void LaunchSmooth(hipStream_t st, const Geo& g, const double* u, double* out) {
  KSmooth<<<Grid(long(g.nx) * g.ny), kBlock, 0, st>>>(g, u, out);
}
void LaunchSmooth2(hipStream_t st, const Geo& g, const double* u, double* out) {
  const unsigned grid = unsigned(((g.nx + kSmTX - 1) / kSmTX) * ((g.ny + kSmTY - 1) / kSmTY));
  KSmooth2<<<grid, kBlock, 0, st>>>(g, u, out);
}
void LaunchPhase0Init(hipStream_t st, const Geo& g, const Phases& ph, const double* pd1,
                      const double* pd2, double* pd0) {
  KPhase0Init<<<Grid(long(g.nx) * g.ny), kBlock, 0, st>>>(g, ph, pd1, pd2, pd0);
}
void LaunchNanCheck(hipStream_t st, const double* a, const double* b, const double* c, int n,
                    int* flag) {
  KNanCheck<<<Grid(n), kBlock, 0, st>>>(a, b, c, n, flag);
}
void LaunchHeatRows(hipStream_t st, const Geo& g, const HeatFaces& hb, const double* const alpha[3],
                    int np, const double cond_i[3], const HeatIn& in_, double* cond, double* gx,
                    double* gy, double* s, double* w, double* c, double* e, double* n,
                    double* k) {
  PhaseCells pc{};
  for (int p = 0; p < np; ++p) {
    pc.alpha[p] = alpha[p];
    pc.density[p] = cond_i[p];
  }
  const dim3 grid = Grid(long(g.nx) * g.ny);
  KHeatGrad<<<grid, kBlock, 0, st>>>(g, hb, pc, np, in_.t, cond, gx, gy);
  HeatIn in = in_;
  in.cond = cond;
  in.gx = gx;
  in.gy = gy;
  KHeatRows<<<grid, kBlock, 0, st>>>(g, hb, in, s, w, c, e, n, k);
}
void LaunchSimpler(hipStream_t st, const Geo& g, Rows vel, const double* k0, const double* k1,
                   const double* uc0, const double* uc1, const double* up0, const double* up1,
                   const double* frx, const double* fry, const double* gpx, const double* gpy,
                   const double* ffex, const double* ffey, const double* ap, const double* flux,
                   double rc, Rows pres, const double* p, int pinned, double pval, double* ev0,
                   double* ev1, double* k) {
  const dim3 grid = Grid(long(g.nx) * g.ny);
  KSimplerCells<<<grid, kBlock, 0, st>>>(g, vel, k0, k1, uc0, uc1, up0, up1, frx, fry, gpx, gpy,
                                         ev0, ev1);
  KSimplerRhs<<<grid, kBlock, 0, st>>>(g, ev0, ev1, ffex, ffey, ap, flux, uc0, uc1, rc, pres, p,
                                       pinned, pval, k);
}
void LaunchGatherInlet(hipStream_t st, const Geo& g, const int kind[4], const double* const pd[3],
                       int np, double* pdb) {
  SideKinds sk{{kind[0], kind[1], kind[2], kind[3]}};
  PhaseCells pc{};
  for (int p = 0; p < np; ++p) pc.alpha[p] = pd[p];
  const int nb = 2 * (g.nx + g.ny);
  KGatherInlet<<<(nb + kBlock - 1) / kBlock, kBlock, 0, st>>>(g, sk, pc, np, pdb);
}
void LaunchOutletUpdate(hipStream_t st, const Geo& g, const int kind[4], const double* u0,
                        const double* u1, double* bv0, double* bv1, double* scratch) {
  SideKinds sk{{kind[0], kind[1], kind[2], kind[3]}};
  const int nb = 2 * (g.nx + g.ny);
  const int grid = (nb + kBlock - 1) / kBlock;
  double *term = scratch, *area = scratch + nb, *corr = scratch + 2 * nb;
  KOutletTerms<<<grid, kBlock, 0, st>>>(g, sk, u0, u1, bv0, bv1, term, area);
  KOutletSum<<<1, kBlock, 0, st>>>(g, sk, term, area, corr);
  KOutletApply<<<grid, kBlock, 0, st>>>(g, sk, corr, bv0, bv1);
}
int BoundaryFluxBlocks(const Geo& g) { return (2 * (g.nx + g.ny) + kBlock - 1) / kBlock; }
void LaunchBoundaryFlux(hipStream_t st, const Geo& g, const double* flux,
                        const double* const alpha[3], int np, const double density[3], double dt,
                        double* part, double* acc) {
  PhaseCells pc{};
  for (int p = 0; p < np; ++p) {
    pc.alpha[p] = alpha[p];
    pc.density[p] = density[p];
  }
  const int nb = BoundaryFluxBlocks(g);
  KBoundaryFluxPartial<<<nb, kBlock, 0, st>>>(g, flux, pc, np, part);
  KBoundaryFluxFinal<<<1, 64, 0, st>>>(part, nb, pc, np, dt, acc);
}
void LaunchAutoDt(hipStream_t st, const Geo& g, const double* flux, double* part, double* out) {
  const int nb = StatsBlocks(g);
  KAutoDtPartial<<<nb, kBlock, 0, st>>>(g, flux, part);
  KAutoDtFinal<<<1, kBlock, 0, st>>>(part, nb, out);
}
int StatsBlocks(const Geo& g) { return int((long(g.nx) * g.ny + kBlock - 1) / kBlock); }
void LaunchStats(hipStream_t st, const Geo& g, const double* xn, const double* yn,
                 const double* alpha, const double* pd, const double* u0, const double* u1,
                 double* part, double* out7) {
  int nb = StatsBlocks(g);
  KStatsPartial<<<nb, kBlock, 0, st>>>(g, xn, yn, alpha, pd, u0, u1, part);
  KStatsFinal<<<1, 64, 0, st>>>(part, nb, out7);
}

}  // namespace hg
