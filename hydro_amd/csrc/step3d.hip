// step3d.hip -- the 3D time step (hydro3d) on MI355X behind hydro_gpu3d.h.
//
// hydro<geom3d::MeshStructured<double>>::step (hydro3d.cpp:17-20 registers
// the module; hydro2d.hpp:1531-1621 is the step, fluid.hpp:814-1158 the
// SIMPLE iteration, conv_diff.hpp:130-251 the momentum rows,
// advection.hpp:417-545 the advection, hydro2d.hpp:1404-1430 the
// properties) on the structured 3D mesh (mesh3d.hpp:240-460).  Every kernel
// is one loop of the reference's algorithm over cells or faces, with each
// value's floating-point operations in the reference's order (the file is
// built with -ffp-contract=off), so the step is bit for bit the reference's
// with the jacobi solvers.  Geometry (face centres, surfaces, areas,
// normals, vectors to cells; cell centres and volumes) is computed on the
// host with mesh3d.hpp's formulas and kept per face and per cell, exact for
// any spacing.  Solvers: jacobi (linear.hpp:750-783: grid sweeps, the max
// update per sweep folded on the device, the reference's stop rule), cg
// (FP64 Jacobi-PCG on the 7-point rows to pressure_tolerance_gpu) and, for
// the pressure, fcg_mg (FP64 flexible CG with an FP32 multigrid V-cycle,
// mg3d.h, to the same tolerance).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hydro_gpu3d.h"
#include "mg3d.h"

#define H3_CHECK(x)                                                                    \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess)                                                              \
      throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e_) + " at " #x); \
  } while (0)

namespace hg3 {
namespace {

constexpr int kB = 256;
inline unsigned Grid(long n) { return unsigned((n + kB - 1) / kB); }

enum BcKind { BC_NONE = 0, BC_VALUE, BC_DERIV, BC_EXTRAP };
struct Bc {
  int kind[6];
  double val[6];
};
Bc MakeBc(BcKind k, double v = 0.) {
  Bc b;
  for (int s = 0; s < 6; ++s) {
    b.kind[s] = k;
    b.val[s] = v;
  }
  return b;
}

// Geometry, device pointers.
struct G {
  int nx, ny, nz;
  long nc, nfx, nfy, nf;
  const double *fc[3], *surf[3], *nrm[3], *to0[3], *to1[3];
  const double *area, *vol;
  const double *cc[3];
};

__device__ __forceinline__ long FaceOf(const G& g, long c, int q) {
  const int i = int(c % g.nx), j = int((c / g.nx) % g.ny), k = int(c / (long(g.nx) * g.ny));
  switch (q) {
    case 0: return i + long(g.nx + 1) * (j + long(g.ny) * k);
    case 1: return i + 1 + long(g.nx + 1) * (j + long(g.ny) * k);
    case 2: return g.nfx + i + long(g.nx) * (j + long(g.ny + 1) * k);
    case 3: return g.nfx + i + long(g.nx) * (j + 1 + long(g.ny + 1) * k);
    case 4: return g.nfx + g.nfy + i + long(g.nx) * (j + long(g.ny) * k);
    default: return g.nfx + g.nfy + i + long(g.nx) * (j + long(g.ny) * (k + 1));
  }
}
__device__ __forceinline__ double Out(int q) { return (q % 2 == 0) ? -1. : 1.; }
// Cells on the minus and plus side of face f (-1 outside the grid), from the
// face number of the structured mesh (x-faces, then y-faces, then z-faces,
// each with x fastest): the same maps as the host's fcm / fcp arrays
// (BuildGeometry), computed instead of loaded -- two 8-byte loads per face
// were most of the face kernels' memory traffic.  32-bit arithmetic:
// hg3_create refuses grids of 2^31 or more faces.
struct FC {
  long m, p;
};
__device__ __forceinline__ FC FaceCells(const G& g, long f) {
  const unsigned nx = unsigned(g.nx), ny = unsigned(g.ny), nz = unsigned(g.nz);
  const unsigned nfx = unsigned(g.nfx), nfy = unsigned(g.nfy);
  const unsigned uf = unsigned(f);
  if (uf < nfx) {
    const unsigned i = uf % (nx + 1u), r = uf / (nx + 1u);  // r = j + ny k
    const long c = long(i) + long(nx) * r;
    return FC{i > 0u ? c - 1 : -1, i < nx ? c : -1};
  }
  if (uf < nfx + nfy) {
    const unsigned q = uf - nfx;  // i + nx (j + (ny + 1) k)
    const unsigned i = q % nx, t = q / nx, j = t % (ny + 1u), k = t / (ny + 1u);
    const long c = long(i) + long(nx) * (long(j) + long(ny) * k);
    return FC{j > 0u ? c - long(nx) : -1, j < ny ? c : -1};
  }
  const unsigned q = uf - nfx - nfy;  // i + nx (j + ny k), k = 0..nz
  const unsigned nxy = nx * ny, k = q / nxy;
  return FC{k > 0u ? long(q) - long(nxy) : -1, k < nz ? long(q) : -1};
}
__device__ __forceinline__ bool Inner(const G& g, long f) {
  const FC fc = FaceCells(g, f);
  return fc.m >= 0 && fc.p >= 0;
}
__device__ __forceinline__ double Dot3(double ax, double ay, double az, double bx, double by,
                                       double bz) {
  double s = 0;
  s += ax * bx;
  s += ay * by;
  s += az * bz;
  return s;
}
__device__ __forceinline__ double Norm3(double x, double y, double z) {
  double s = 0;
  s += x * x;
  s += y * y;
  s += z * z;
  return sqrt(s);
}
__device__ __forceinline__ double OutS(const G& g, long f, int q, int d) {
  return g.surf[d][f] * Out(q);
}
__device__ __forceinline__ double ToDist(const G& g, long f, int id) {
  const double* const* t = id == 0 ? g.to0 : g.to1;
  return Norm3(t[0][f], t[1][f], t[2][f]);
}
// boundary side of face f (0..5), the face being a boundary face
__device__ __forceinline__ int SideOf(const G& g, long f) {
  if (f < g.nfx) return (f % (g.nx + 1)) == 0 ? 0 : 1;
  if (f < g.nfx + g.nfy) {
    const long r = f - g.nfx;
    return ((r / g.nx) % (g.ny + 1)) == 0 ? 2 : 3;
  }
  const long r = f - g.nfx - g.nfy;
  return (r / (long(g.nx) * g.ny)) == 0 ? 4 : 5;
}

// ---- interpolation (solver.hpp:392-470)
// Inner faces: the mean of the two cells.  Boundary faces: with bc_here,
// the value and zero-derivative conditions (which read only the face's own
// cell, as in KInterpBoundary), else 0 for KInterpBoundary to fill.
__global__ void KInterpInner(G g, const double* __restrict__ u, double* __restrict__ res, Bc bc,
                             int bc_here) {
  const long f = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (f >= g.nf) return;
  const FC fc = FaceCells(g, f);
  double v = 0.;
  if (fc.m >= 0 && fc.p >= 0) {
    v = u[fc.m] * (1. - 0.5) + u[fc.p] * 0.5;
  } else if (bc_here) {
    const int side = SideOf(g, f);
    const int kind = bc.kind[side];
    const int id = fc.m >= 0 ? 0 : 1;
    const long c = id == 0 ? fc.m : fc.p;
    const double factor = (id == 0 ? 1. : -1.);
    if (kind == BC_VALUE) {
      v = bc.val[side];
    } else if (kind == BC_DERIV) {
      const double alpha = ToDist(g, f, id) * factor;
      v = u[c] + bc.val[side] * alpha;
    }
  }
  res[f] = v;
}
// boundary faces of one direction (x: sides 0,1; y: 2,3; z: 4,5); the
// directions run in face order, so an extrapolated face sees the values of
// earlier directions' boundary faces as the reference's sequential loop does
__global__ void KInterpBoundary(G g, int dir, const double* __restrict__ u, double* res, Bc bc) {
  const long t = blockIdx.x * long(blockDim.x) + threadIdx.x;
  long f;
  if (dir == 0) {
    const long per = long(g.ny) * g.nz;
    if (t >= 2 * per) return;
    const int i = t < per ? 0 : g.nx;
    const long r = t % per;
    f = i + long(g.nx + 1) * r;
  } else if (dir == 1) {
    const long per = long(g.nx) * g.nz;
    if (t >= 2 * per) return;
    const int j = t < per ? 0 : g.ny;
    const long r = t % per, i = r % g.nx, k = r / g.nx;
    f = g.nfx + i + long(g.nx) * (j + long(g.ny + 1) * k);
  } else {
    const long per = long(g.nx) * g.ny;
    if (t >= 2 * per) return;
    const int k = t < per ? 0 : g.nz;
    const long r = t % per;
    f = g.nfx + g.nfy + r + per * k;
  }
  const int side = SideOf(g, f);
  const int kind = bc.kind[side];
  if (kind == BC_NONE) return;
  const int id = FaceCells(g, f).m >= 0 ? 0 : 1;
  const long c = id == 0 ? FaceCells(g, f).m : FaceCells(g, f).p;
  const double factor = (id == 0 ? 1. : -1.);
  if (kind == BC_VALUE) {
    res[f] = bc.val[side];
  } else if (kind == BC_DERIV) {
    const double alpha = ToDist(g, f, id) * factor;
    res[f] = u[c] + bc.val[side] * alpha;
  } else {
    const double nx_ = g.nrm[0][f] * factor, ny_ = g.nrm[1][f] * factor, nz_ = g.nrm[2][f] * factor;
    const double dist = ToDist(g, f, id);
    double nom = u[c] / dist;
    double den = 1. / dist;
    const double volume = g.vol[c];
    for (int q = 0; q < 6; ++q) {
      const long nf_ = FaceOf(g, c, q);
      const double dn = Dot3(OutS(g, nf_, q, 0), OutS(g, nf_, q, 1), OutS(g, nf_, q, 2), nx_, ny_, nz_);
      if (nf_ == f)
        den -= dn / volume;
      else
        nom += res[nf_] * dn / volume;
    }
    res[f] = nom / den;
  }
}

// ---- gradient (solver.hpp:658-677), average (solver.hpp:621-634)
__global__ void KGradient(G g, const double* __restrict__ uf, double* __restrict__ gx,
                          double* __restrict__ gy, double* __restrict__ gz) {
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c >= g.nc) return;
  double sx = 0, sy = 0, sz = 0;
  for (int q = 0; q < 6; ++q) {
    const long f = FaceOf(g, c, q);
    const double u = uf[f];
    sx += OutS(g, f, q, 0) * u;
    sy += OutS(g, f, q, 1) * u;
    sz += OutS(g, f, q, 2) * u;
  }
  gx[c] = sx / g.vol[c];
  gy[c] = sy / g.vol[c];
  gz[c] = sz / g.vol[c];
}
__global__ void KAverage(G g, const double* __restrict__ uf, double* __restrict__ u) {
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c >= g.nc) return;
  double s = 0;
  for (int q = 0; q < 6; ++q) s += uf[FaceOf(g, c, q)];
  u[c] = s / 6.;
}

// ---- restored force (fluid.hpp:602-631)
__global__ void KRestoredForce(G g, const double* __restrict__ fex, const double* __restrict__ fey,
                               const double* __restrict__ fez, double* __restrict__ frx,
                               double* __restrict__ fry, double* __restrict__ frz) {
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c >= g.nc) return;
  double sx = 0, sy = 0, sz = 0;
  for (int q = 0; q < 6; ++q) {
    const long f = FaceOf(g, c, q);
    const double dn = Dot3(fex[f], fey[f], fez[f], g.nrm[0][f], g.nrm[1][f], g.nrm[2][f]);
    const double dist = Norm3(g.fc[0][f] - g.cc[0][c], g.fc[1][f] - g.cc[1][c], g.fc[2][f] - g.cc[2][c]);
    sx += g.surf[0][f] * dn * dist;
    sy += g.surf[1][f] * dn * dist;
    sz += g.surf[2][f] * dn * dist;
  }
  frx[c] = sx / g.vol[c];
  fry[c] = sy / g.vol[c];
  frz[c] = sz / g.vol[c];
}

// ---- explicit viscosity of component d (fluid.hpp:835-854): fo += sum
// over faces of the face gradient times mu_f S_out[d], / V
__global__ void KViscosity(G g, int d, const double* __restrict__ ffmu, const double* __restrict__ gfx,
                           const double* __restrict__ gfy, const double* __restrict__ gfz,
                           double* fox, double* foy, double* foz) {
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c >= g.nc) return;
  double sx = 0, sy = 0, sz = 0;
  for (int q = 0; q < 6; ++q) {
    const long f = FaceOf(g, c, q);
    const double w = ffmu[f] * OutS(g, f, q, d);
    sx += gfx[f] * w;
    sy += gfy[f] * w;
    sz += gfz[f] * w;
  }
  fox[c] += sx / g.vol[c];
  foy[c] += sy / g.vol[c];
  foz[c] += sz / g.vol[c];
}
// force assembly (fluid.hpp:857-870), zero sources: fo += (-grad p + restored
// force + 0) + u * (rho * 0 - 0)
struct Force3 {
  const double* gp[3];
  const double* fr[3];
  const double* u[3];
  double* fo[3];
};
__global__ void KForceAssembly(G g, const double* __restrict__ rho_s, Force3 p) {
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c >= g.nc) return;
  const double s = rho_s[c] * 0. - 0.;
  for (int d = 0; d < 3; ++d) p.fo[d][c] += ((p.gp[d][c] * (-1.) + p.fr[d][c]) + 0.) + p.u[d][c] * s;
}
// initial flux (fluid.hpp:770-784): u_f . S - meshvel . S (meshvel 0)
struct Faces3c {
  const double *x, *y, *z;
};
__global__ void KFaceFlux(G g, Faces3c fu, double* __restrict__ fl) {
  const long f = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (f >= g.nf) return;
  const double sx = g.surf[0][f], sy = g.surf[1][f], sz = g.surf[2][f];
  fl[f] = Dot3(fu.x[f], fu.y[f], fu.z[f], sx, sy, sz) - Dot3(0., 0., 0., sx, sy, sz);
}

// ---- momentum face expressions of one component (conv_diff.hpp:150-186;
// solver.hpp:201-243 upwind, 283-341 diffusion)
struct FaceExpr {
  double *cam, *cap, *cak, *dcm, *dcp, *dck;
};
__global__ void KMomentumFaces(G g, const double* __restrict__ probe, const double* __restrict__ ffmu,
                               const double* __restrict__ gx, const double* __restrict__ gy,
                               const double* __restrict__ gz, Bc vbc, FaceExpr e) {
  const long f = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (f >= g.nf) return;
  const double F = probe[f], mu = ffmu[f], A_ = g.area[f];
  if (Inner(g, f)) {
    const long cm = FaceCells(g, f).m, cp = FaceCells(g, f).p;
    double am, ap, ak;
    if (F > 1e-10) {
      am = 1.;
      ap = 0.;
      ak = -Dot3(gx[cm], gy[cm], gz[cm], g.to0[0][f], g.to0[1][f], g.to0[2][f]);
    } else if (F < -1e-10) {
      am = 0.;
      ap = 1.;
      ak = -Dot3(gx[cp], gy[cp], gz[cp], g.to1[0][f], g.to1[1][f], g.to1[2][f]);
    } else {
      am = 0.5;
      ap = 0.5;
      ak = 0.;
    }
    e.cam[f] = am * F;
    e.cap[f] = ap * F;
    e.cak[f] = ak * F;
    const double al = 1. / Norm3(g.to1[0][f] - g.to0[0][f], g.to1[1][f] - g.to0[1][f],
                                 g.to1[2][f] - g.to0[2][f]);
    e.dcm[f] = ((-al) * (-mu)) * A_;
    e.dcp[f] = (al * (-mu)) * A_;
    e.dck[f] = (0. * (-mu)) * A_;
  } else {
    const int id = FaceCells(g, f).m >= 0 ? 0 : 1;
    const double val = vbc.val[SideOf(g, f)];
    e.cam[f] = e.cap[f] = 0.;
    e.cak[f] = val * F;
    const double factor = (id == 0 ? 1. : -1.);
    const double al = 1. / ToDist(g, f, id) * factor;
    const double dcc = ((-al) * (-mu)) * A_;
    e.dcm[f] = id == 0 ? dcc : 0.;
    e.dcp[f] = id == 0 ? 0. : dcc;
    e.dck[f] = ((al * val) * (-mu)) * A_;
  }
}

// row term of face q (W, E, S, N, B, F) in cell-index order B S W C E N F
__device__ __constant__ int kFaceTerm[6] = {2, 4, 1, 5, 0, 6};
__device__ __forceinline__ bool Has(const G& g, long c, int q) {
  const int i = int(c % g.nx), j = int((c / g.nx) % g.ny), k = int(c / (long(g.nx) * g.ny));
  switch (q) {
    case 0: return k > 0;
    case 1: return j > 0;
    case 2: return i > 0;
    case 3: return true;
    case 4: return i < g.nx - 1;
    case 5: return j < g.ny - 1;
    default: return k < g.nz - 1;
  }
}
__device__ __forceinline__ long Off(const G& g, int q) {
  const long nxy = long(g.nx) * g.ny;
  switch (q) {
    case 0: return -nxy;
    case 1: return -long(g.nx);
    case 2: return -1;
    case 3: return 0;
    case 4: return 1;
    case 5: return long(g.nx);
    default: return nxy;
  }
}
struct Rows {
  double* t[7];
  double* k;
};
struct MomCell {
  const double *rho_s, *utp, *utc, *fo, *up;
  double co0, co1, co2, relax;
};
__global__ void KMomentumRows(G g, FaceExpr e, MomCell m, int write_coefs, Rows A, double* kk) {
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c >= g.nc) return;
  double Cc = 0., Ck = 0., Dc = 0., Dk = 0.;
  double Cnb[6] = {0, 0, 0, 0, 0, 0}, Dnb[6] = {0, 0, 0, 0, 0, 0};
  for (int q = 0; q < 6; ++q) {
    const long f = FaceOf(g, c, q);
    const double o_ = Out(q);
    const bool self_is_m = (q % 2 == 1);
    const double cself = self_is_m ? e.cam[f] : e.cap[f];
    const double cnb = self_is_m ? e.cap[f] : e.cam[f];
    const double dself = self_is_m ? e.dcm[f] : e.dcp[f];
    const double dnb = self_is_m ? e.dcp[f] : e.dcm[f];
    if (Inner(g, f)) {
      Cc += cself * o_;
      Cnb[q] = cnb * o_;
      Dnb[q] = dnb * o_;
    }
    Dc += dself * o_;
    Ck += e.cak[f] * o_;
    Dk += e.dck[f] * o_;
  }
  const double V = g.vol[c], rho = m.rho_s[c];
  const double cen = (Dc / V) + ((Cc / V + m.co2) * rho);
  double row[7];
  for (int q = 0; q < 6; ++q) row[kFaceTerm[q]] = (Dnb[q] / V) + ((Cnb[q] / V) * rho);
  row[3] = cen;
  const double unst = m.co0 * m.utp[c] + m.co1 * m.utc[c];
  const double kon = ((Dk / V) + ((Ck / V + unst) * rho)) + (-m.fo[c]);
  double ev = kon;
  for (int q = 0; q < 7; ++q)
    if (Has(g, c, q)) ev += m.up[c + Off(g, q)] * row[q];
  kk[c] = ev;
  if (write_coefs)
    for (int q = 0; q < 7; ++q) A.t[q][c] = (q == 3) ? cen / m.relax : (Has(g, c, q) ? row[q] : 0.);
}

// ---- Jacobi (linear.hpp:750-783): one sweep x -> xn with the max |corr|
// folded into diff[s] (non-negative doubles order as their bit patterns).
// The stop rule after sweep s (state = {done, iters, buffer holding the
// result}): converged (iters = s) or over the limit (iters = s + 1).
__device__ __forceinline__ bool JacobiStop(const unsigned long long* diff, int s, double tol,
                                           int limit, int& iters) {
  const double d = __longlong_as_double((long long)diff[s]);
  if (!(d > tol)) {
    iters = s;
    return true;
  }
  if (!(s < limit)) {
    iters = s + 1;
    return true;
  }
  return false;
}
// Sweep s first applies the stop rule of sweep s - 1 (complete: its kernel
// has finished), so no separate check runs between sweeps: every block
// takes the same decision from diff[s - 1], and block 0 records it.
__global__ void KJacobiSweep(G g, Rows A, const double* __restrict__ k, const double* __restrict__ x,
                             double* __restrict__ xn, double omega, unsigned long long* diff,
                             int s, int* state, double tol, int limit) {
  if (state[0]) return;  // decided after sweep s - 2 or earlier
  int iters = 0;
  if (s > 0 && JacobiStop(diff, s - 1, tol, limit, iters)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      state[0] = 1;
      state[1] = iters;
      state[2] = s & 1;  // sweep s - 1 wrote buffer s & 1
    }
    return;
  }
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  double m = 0.;
  if (c < g.nc) {
    double sum = 0.;
    for (int q = 0; q < 7; ++q)
      if (q != 3 && Has(g, c, q)) sum += A.t[q][c] * x[c + Off(g, q)];
    const double value = -(k[c] + sum) / A.t[3][c];
    const double corr = value - x[c];
    m = fabs(corr);
    xn[c] = x[c] + corr * omega;
  }
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0 && m > 0.)
    atomicMax(&diff[s], (unsigned long long)__double_as_longlong(m));
}
// the stop rule after the last sweep s of a batch
__global__ void KJacobiCheck(const unsigned long long* diff, int s, double tol, int limit,
                             int* state) {
  if (state[0]) return;
  int iters = 0;
  if (JacobiStop(diff, s, tol, limit, iters)) {
    state[0] = 1;
    state[1] = iters;
    state[2] = (s + 1) & 1;  // sweep s wrote buffer (s + 1) & 1
  }
}

// The three velocity components' Jacobi solves (one matrix, three right-hand
// sides, each from zero with its own stop rule) in one kernel per sweep:
// the seven row entries of a cell are loaded once for the components still
// running, and each component's update is the single-component sweep's,
// operation for operation.  Component d keeps its max |corr| in
// diff[d * stride + s] and its state (done, iters, buffer) at state[4 d].
struct Jac3 {
  const double* k[3];
  const double* x[3];  // sweep s's input iterate, per component
  double* xn[3];       // its output
};
__global__ void KJacobiSweep3(G g, Rows A, Jac3 J, double omega, unsigned long long* diff,
                              int stride, int s, int* state, double tol, int limit) {
  bool act[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    act[d] = !state[4 * d];  // decided after sweep s - 2 or earlier: off
    int iters = 0;
    if (act[d] && s > 0 && JacobiStop(diff + long(d) * stride, s - 1, tol, limit, iters)) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        state[4 * d] = 1;
        state[4 * d + 1] = iters;
        state[4 * d + 2] = s & 1;  // sweep s - 1 wrote buffer s & 1
      }
      act[d] = false;
    }
  }
  if (!act[0] && !act[1] && !act[2]) return;  // uniform over the grid
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  double m[3] = {0., 0., 0.};
  if (c < g.nc) {
    bool has[7];
    double t[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      has[q] = q != 3 && Has(g, c, q);
      t[q] = (q == 3 || has[q]) ? A.t[q][c] : 0.;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (!act[d]) continue;
      const double* x = J.x[d];
      double sum = 0.;
#pragma unroll
      for (int q = 0; q < 7; ++q)
        if (has[q]) sum += t[q] * x[c + Off(g, q)];
      const double value = -(J.k[d][c] + sum) / t[3];
      const double corr = value - x[c];
      m[d] = fabs(corr);
      J.xn[d][c] = x[c] + corr * omega;
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (!act[d]) continue;
    double v = m[d];
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0 && v > 0.)
      atomicMax(&diff[long(d) * stride + s], (unsigned long long)__double_as_longlong(v));
  }
}
// the stop rule after the last sweep s of a batch, for each component
__global__ void KJacobiCheck3(const unsigned long long* diff, int stride, int s, double tol,
                              int limit, int* state) {
  const int d = threadIdx.x;
  if (d >= 3 || state[4 * d]) return;
  int iters = 0;
  if (JacobiStop(diff + long(d) * stride, s, tol, limit, iters)) {
    state[4 * d] = 1;
    state[4 * d + 1] = iters;
    state[4 * d + 2] = (s + 1) & 1;  // sweep s wrote buffer (s + 1) & 1
  }
}

// ---- cg: FP64 Jacobi-PCG, partial sums per block folded by one block
__device__ __forceinline__ double BlockSum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.;
  if (threadIdx.x == 0)
    for (int w = 0; w < kB / 64; ++w) t += red[w];
  return t;
}
// grid-stride loops over cells: a grid of at most kCgBlocks blocks leaves
// few partial sums for the one-block folds
constexpr int kCgBlocks = 1024;
#define H3_CELLS(c, n) \
  for (long c = blockIdx.x * long(blockDim.x) + threadIdx.x; c < (n); c += long(gridDim.x) * blockDim.x)
__global__ void KApplyRows(G g, Rows A, const double* __restrict__ x, double* __restrict__ y,
                           const double* __restrict__ p2, double* part) {
  __shared__ double red[kB / 64];
  double v = 0.;
  H3_CELLS(c, g.nc) {
    double s = 0.;
    for (int q = 0; q < 7; ++q)
      if (Has(g, c, q)) s += A.t[q][c] * x[c + Off(g, q)];
    y[c] = s;
    if (p2) v += p2[c] * s;
  }
  if (part) {
    const double t = BlockSum(v, red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
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
__global__ void KFold(const double* part, int n, double* out) {
  __shared__ double red[kB / 64];
  const double v = StridedSum<4>(part, n, threadIdx.x, blockDim.x);  // n <= kCgBlocks = 4 kB
  const double t = BlockSum(v, red);
  if (threadIdx.x == 0) *out = t;
}
// elementwise helpers of the PCG
__global__ void KCgStart(long n, const double* k, const double* mean, const double* diag,
                         double* b, double* r, double* z, double* p, double* x, double* part) {
  __shared__ double red[kB / 64];
  double v = 0.;
  H3_CELLS(c, n) {
    const double bv = -k[c] - *mean;
    b[c] = bv;
    r[c] = bv;
    const double zv = bv / diag[c];
    z[c] = zv;
    p[c] = zv;
    x[c] = 0.;
    v += bv * zv;
  }
  const double t = BlockSum(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}
__global__ void KSumK(long n, const double* k, double* part) {
  __shared__ double red[kB / 64];
  double v = 0.;
  H3_CELLS(c, n) v += -k[c];
  const double t = BlockSum(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}
__global__ void KDivide(double* v, double n) { *v = *v / n; }
__global__ void KFormB(long n, const double* k, const double* mean, double* b) {
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c < n) b[c] = -k[c] - *mean;
}
struct CgScal {
  double mean, bb, rz, pq, rr, rz2;
  double converged;  // 1 when the stop was by the tolerance (set with done)
  int done, iters;
};
__global__ void KCgUpdate(long n, const double* p, const double* q, double* x, double* r,
                          const double* diag, double* z, CgScal* sc, double* part_rr,
                          double* part_rz) {
  __shared__ double red[kB / 64];
  if (sc->done) return;
  const double alpha = sc->rz / sc->pq;
  double rr = 0., rz = 0.;
  H3_CELLS(c, n) {
    x[c] += alpha * p[c];
    const double rv = r[c] - alpha * q[c];
    r[c] = rv;
    const double zv = rv / diag[c];
    z[c] = zv;
    rr += rv * rv;
    rz += rv * zv;
  }
  const double t1 = BlockSum(rr, red);
  __syncthreads();
  const double t2 = BlockSum(rz, red);
  if (threadIdx.x == 0) {
    part_rr[blockIdx.x] = t1;
    part_rz[blockIdx.x] = t2;
  }
}
// The stop test; converged records the reason at the stop, since the
// iterations still queued behind it refold partial buffers into rr.
__global__ void KCgCheck(CgScal* sc, double tol2, int max_iters) {
  if (sc->done) return;
  sc->iters += 1;
  const bool conv = sc->rr <= tol2 * sc->bb;
  if (conv || sc->iters >= max_iters) {
    sc->done = 1;
    sc->converged = conv ? 1. : 0.;
  }
}
__global__ void KCgDir(long n, const double* z, double* p, const CgScal* sc) {
  if (sc->done) return;
  const double beta = sc->rz2 / sc->rz;
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c < n) p[c] = z[c] + beta * p[c];
}
__global__ void KCgShift(CgScal* sc) {
  if (sc->done) return;
  sc->rz = sc->rz2;
}
__global__ void KCgResidual(long n, const double* b, const double* ax, double* part) {
  __shared__ double red[kB / 64];
  double v = 0.;
  H3_CELLS(c, n) v += (b[c] - ax[c]) * (b[c] - ax[c]);
  const double t = BlockSum(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// ---- pieces of the iteration
__global__ void KAddTo(long n, const double* a, const double* b, double* out) {
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c < n) out[c] = a[c] + b[c];
}
__global__ void KCoeffMean(G g, Rows A, double* ap) {
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c >= g.nc) return;
  double cs = 0.;
  for (int q = 0; q < 7; ++q)
    if (Has(g, c, q)) cs += A.t[q][c];
  double sum = 0.;
  sum += cs;
  sum += cs;
  sum += cs;
  ap[c] = sum / 3.;
}
struct Faces3 {
  const double *x, *y, *z;
};
__global__ void KRhieChow(G g, Faces3 fu, Faces3 ffgp, Faces3 ffr, Faces3 ffe,
                          const double* __restrict__ p, const double* __restrict__ ffd, double rc,
                          double* __restrict__ fstar, double* __restrict__ fcoef) {
  const long f = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (f >= g.nf) return;
  const double sx = g.surf[0][f], sy = g.surf[1][f], sz = g.surf[2][f];
  const double vfi = Dot3(fu.x[f], fu.y[f], fu.z[f], sx, sy, sz);
  double fs = vfi, fc = 0.;
  if (Inner(g, f)) {
    const long cm = FaceCells(g, f).m, cp = FaceCells(g, f).p;
    const double dn = Norm3(g.to1[0][f] - g.to0[0][f], g.to1[1][f] - g.to0[1][f],
                            g.to1[2][f] - g.to0[2][f]);
    const double wide = Dot3(ffgp.x[f] - ffr.x[f], ffgp.y[f] - ffr.y[f], ffgp.z[f] - ffr.z[f],
                             sx, sy, sz);
    const double compact = (p[cp] - p[cm]) / dn * g.area[f] - Dot3(ffe.x[f], ffe.y[f], ffe.z[f], sx, sy, sz);
    fs = vfi + rc * (wide - compact) / ffd[f] + 0;
    fc = -g.area[f] / (dn * ffd[f]);
  }
  fstar[f] = fs - Dot3(0., 0., 0., sx, sy, sz);
  fcoef[f] = fc;
}
__global__ void KPressureRows(G g, const double* __restrict__ fstar, const double* __restrict__ fcoef,
                              Rows P) {
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c >= g.nc) return;
  double cen = 0., kon = 0.;
  double nb[6] = {0, 0, 0, 0, 0, 0};
  for (int q = 0; q < 6; ++q) {
    const long f = FaceOf(g, c, q);
    const double o_ = Out(q);
    const bool self_is_m = (q % 2 == 1);
    if (Inner(g, f)) {
      const double cself = self_is_m ? -fcoef[f] : fcoef[f];
      const double cnb = self_is_m ? fcoef[f] : -fcoef[f];
      cen += cself * o_;
      nb[q] = cnb * o_;
    }
    kon += fstar[f] * o_;
  }
  kon += -(0. * g.vol[c]);
  for (int q = 0; q < 6; ++q) P.t[kFaceTerm[q]][c] = nb[q];
  P.t[3][c] = cen;
  P.k[c] = kon;
}
__global__ void KCorrectPressure(long n, const double* p_ip, const double* pc, double relax,
                                 double* p_ic) {
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c < n) p_ic[c] = p_ip[c] + relax * pc[c];
}
__global__ void KCorrectVelocity(long n, const double* gx, const double* gy, const double* gz,
                                 const double* ap, double* u0, double* u1, double* u2) {
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c >= n) return;
  u0[c] += gx[c] / (-ap[c]);
  u1[c] += gy[c] / (-ap[c]);
  u2[c] += gz[c] / (-ap[c]);
}
__global__ void KCorrectFlux(G g, const double* fstar, const double* fcoef, const double* pc,
                             double* fl) {
  const long f = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (f >= g.nf) return;
  double v = fstar[f];
  if (Inner(g, f)) {
    v += pc[FaceCells(g, f).m] * (-fcoef[f]);
    v += pc[FaceCells(g, f).p] * fcoef[f];
  }
  fl[f] = v;
}
// convergence indicator: max |u_ip - u_ic| per block
__global__ void KDiffMax(long n, const double* a0, const double* a1, const double* a2,
                         const double* b0, const double* b1, const double* b2, double* part) {
  __shared__ double red[kB / 64];
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  double v = 0.;
  if (c < n) v = Norm3(b0[c] - a0[c], b1[c] - a1[c], b2[c] - a2[c]);
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.;
    for (int w = 0; w < kB / 64; ++w) t = fmax(t, red[w]);
    part[blockIdx.x] = t;
  }
}
__global__ void KMaxFold(const double* part, int n, double* out) {
  __shared__ double red[kB / 64];
  // a thread's first 32 partials loaded at once (indices clamped into the
  // array, selected afterwards; a maximum is exact in any order), the rest
  // in a loop
  constexpr int K = 32;
  const int t = threadIdx.x, last = n > 0 ? n - 1 : 0;
  double w[K];
#pragma unroll
  for (int k = 0; k < K; ++k) w[k] = part[min(t + k * kB, last)];
  double v = 0.;
#pragma unroll
  for (int k = 0; k < K; ++k)
    if (t + k * kB < n) v = fmax(v, w[k]);
  for (int i = t + K * kB; i < n; i += kB) v = fmax(v, part[i]);
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.;
    for (int w = 0; w < kB / 64; ++w) t = fmax(t, red[w]);
    *out = t;
  }
}
__global__ void KExtrapolate(long n, const double* tc, const double* tp, double ge, double* ic) {
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c < n) ic[c] = tc[c] + (tc[c] - tp[c]) * ge;
}
__global__ void KNan(long n, const double* a, int bit, int* flag) {
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c < n && !(a[c] * 0. == 0.)) atomicOr(flag, bit);
}

// ---- advection (advection.hpp:440-539)
__device__ __forceinline__ double Superbee(double p, double q) {
  if (p > 0. && q > 0.) return fmax(fmin(2 * p, q), fmin(p, 2 * q));
  if (p < 0. && q < 0.) return -fmax(fmin(-2 * p, -q), fmin(-p, -2 * q));
  return 0.;
}
__global__ void KAdvFaces(G g, const double* __restrict__ u, const double* __restrict__ gx,
                          const double* __restrict__ gy, const double* __restrict__ gz,
                          const double* __restrict__ vf, double* __restrict__ fu) {
  const long f = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (f >= g.nf) return;
  if (Inner(g, f)) {
    const long P = FaceCells(g, f).m, E = FaceCells(g, f).p;
    const double v = vf[f] + 0.;
    if (v > 1e-8) {
      fu[f] = u[P] + 0.5 * Superbee(u[E] - u[P], -4. * Dot3(gx[P], gy[P], gz[P], g.to0[0][f],
                                                             g.to0[1][f], g.to0[2][f]) -
                                                      (u[E] - u[P]));
    } else if (v < -1e-8) {
      fu[f] = u[E] - 0.5 * Superbee(u[E] - u[P], 4. * Dot3(gx[E], gy[E], gz[E], g.to1[0][f],
                                                            g.to1[1][f], g.to1[2][f]) -
                                                     (u[E] - u[P]));
    } else {
      fu[f] = 0.5 * (u[P] + u[E]);
    }
  } else {
    const int id = FaceCells(g, f).m >= 0 ? 0 : 1;
    const long cc = id == 0 ? FaceCells(g, f).m : FaceCells(g, f).p;
    const double factor = (id == 0 ? 1. : -1.);
    const double al = ToDist(g, f, id) * factor;
    fu[f] = u[cc] + 0. * al;
  }
}
__global__ void KAdvUpdate(G g, const double* __restrict__ fu, const double* __restrict__ vf,
                           double dt, double* u) {
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c >= g.nc) return;
  double fs = 0.;
  for (int q = 0; q < 6; ++q) {
    const long f = FaceOf(g, c, q);
    fs += fu[f] * (vf[f] + 0.) * Out(q);
  }
  double v = u[c];
  v += -dt / g.vol[c] * fs;
  v += dt * 0. * 0.;
  double sh = 0.;
  for (int q = 0; q < 6; ++q) sh += Out(q) * 0.;
  v += dt * sh / g.vol[c];
  u[c] = v;
}

// ---- properties (hydro2d.hpp:1404-1430)
struct Phases3 {
  int np;
  double density[2], viscosity[2];
  const double* pd[2];
  double* alpha[2];
};
__global__ void KProperties(long n, Phases3 ph, double* rho, double* mu) {
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c >= n) return;
  double a[2], sum = 0.;
  for (int i = 0; i < ph.np; ++i) {
    a[i] = ph.pd[i][c] / ph.density[i];
    sum += a[i];
  }
  for (int i = 0; i < ph.np; ++i) {
    a[i] /= sum;
    ph.alpha[i][c] = a[i];
  }
  double r = 0., m = 0.;
  for (int i = 0; i < ph.np; ++i) r += ph.density[i] * a[i];
  for (int i = 0; i < ph.np; ++i) m += ph.viscosity[i] * a[i];
  rho[c] = r;
  mu[c] = m;
}
__global__ void KForce(long n, const double* rho, double gd, double fd, double* f) {
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c < n) f[c] = gd * rho[c] + fd;
}

// BDF coefficients (DerivCoeffs of the reference's time derivative)
std::vector<double> DerivCoeffs(double target, const std::vector<double>& args) {
  const size_t size = args.size();
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

}  // namespace
}  // namespace hg3

using namespace hg3;

struct hg3_ctx {
  hg3_config cfg3;
  hg_config cfg;
  hipStream_t st = nullptr;
  std::vector<void*> allocs;
  std::string err;
  G g{};
  int np = 1;
  long nc = 0, nf = 0;
  double dt = 0., dt_adv = 0., t_fluid = 0., t_adv = 0.;
  double co[3] = {0, 0, 0};
  double grav[3] = {0, 0, 0}, force[3] = {0, 0, 0};
  // fields (time_curr, time_prev, iter_curr, iter_prev)
  double *u_tc[3], *u_tp[3], *u_ic[3], *u_ip[3];
  double *p_tc, *p_tp, *p_ic, *p_ip;
  double *fl_tc, *fl_tp, *fl_ic, *fl_ip;
  double *pd[2], *alpha[2];
  double *rho, *rho_s, *mu, *mu_s, *f[3];
  // work
  double *ffe[3], *fr[3], *ffr[3], *ffmu, *ffp, *gp[3], *ffgp[3], *fo[3];
  double *ff, *g3[3], *gf3[3];
  FaceExpr fe{};
  Rows A{}, P{};
  double* kk[3];
  double *corr, *corr2, *a_p, *ffd, *fu[3], *fstar, *fcoef, *pc, *ffpc;
  double* jbuf[3][2];  // the velocity components' Jacobi iterates (jacobi velocity solver)
  double *part, *scal_d;
  unsigned long long* jdiff;
  int* jstate;
  int* flag;
  // cg
  double *cb, *cr, *cz, *cp, *cq, *cx;
  CgScal* cs;
  CgScal* cs_h;
  int jlimit = 0;
  bool p_converged = true;  // the last pressure solve met its tolerance (cg, fcg_mg)
  int fluid_iter = 0, cd_iter = 0;
  double last_ind = 1.;
  int* state_h = nullptr;
  double* scal_h = nullptr;

  double* Alloc(long n) {
    void* p = nullptr;
    H3_CHECK(hipMalloc(&p, size_t(std::max<long>(n, 1)) * sizeof(double)));
    H3_CHECK(hipMemsetAsync(p, 0, size_t(std::max<long>(n, 1)) * sizeof(double), st));
    allocs.push_back(p);
    return (double*)p;
  }
  std::unique_ptr<hg3::Mg3Fcg> mg;  // fcg_mg pressure solver
  ~hg3_ctx() {
    if (st) (void)hipStreamSynchronize(st);
    mg.reset();
    for (void* p : allocs) (void)hipFree(p);
    if (state_h) (void)hipHostFree(state_h);
    if (scal_h) (void)hipHostFree(scal_h);
    if (cs_h) (void)hipHostFree(cs_h);
    if (st) (void)hipStreamDestroy(st);
  }
  void Sync() { H3_CHECK(hipStreamSynchronize(st)); }
};

namespace {

thread_local std::string g_create_error3;

struct Err3 : std::runtime_error {
  hg_status code;
  Err3(hg_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

hg_status Guard3(hg3_ctx* c, const std::function<void()>& fn);

// geometry on the host with mesh3d.hpp's formulas (the oracle's), uploaded
void BuildGeometry(hg3_ctx* c) {
  const hg3_config& k = c->cfg3;
  const int nx = k.base.nx, ny = k.base.ny, nz = k.nz;
  const double a[3] = {k.base.a[0], k.base.a[1], k.a_z}, b[3] = {k.base.b[0], k.base.b[1], k.b_z};
  const int nn[3] = {nx, ny, nz};
  const long nc = long(nx) * ny * nz;
  const long nfx = long(nx + 1) * ny * nz, nfy = long(nx) * (ny + 1) * nz;
  const long nf = nfx + nfy + long(nx) * ny * (nz + 1);
  struct P3 {
    double v[3];
  };
  auto node = [&](int i, int j, int kk) {
    const int m[3] = {i, j, kk};
    P3 r;
    for (int d = 0; d < 3; ++d) r.v[d] = a[d] + (double(m[d]) / double(nn[d])) * (b[d] - a[d]);
    return r;
  };
  std::vector<double> fc[3], surf[3], nrm[3], t0[3], t1[3], area(nf), vol(nc), cc[3];
  for (int d = 0; d < 3; ++d) {
    fc[d].assign(nf, 0.);
    surf[d].assign(nf, 0.);
    nrm[d].assign(nf, 0.);
    t0[d].assign(nf, 0.);
    t1[d].assign(nf, 0.);
    cc[d].assign(nc, 0.);
  }
  std::vector<long> fcm(nf), fcp(nf);
  auto cell = [&](int i, int j, int kk) { return long(i) + long(nx) * (j + long(ny) * kk); };
  auto face_geom = [&](long f, const P3 (&nd)[4]) {
    double s3[3] = {0, 0, 0};
    for (int q = 0; q < 4; ++q)
      for (int d = 0; d < 3; ++d) s3[d] += nd[q].v[d];
    for (int d = 0; d < 3; ++d) fc[d][f] = s3[d] / 4.;
    double sf[3] = {0, 0, 0};
    for (int q = 2; q < 4; ++q) {
      double u[3], w[3];
      for (int d = 0; d < 3; ++d) {
        u[d] = nd[q - 1].v[d] - nd[0].v[d];
        w[d] = nd[q].v[d] - nd[0].v[d];
      }
      const double cr[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2],
                            u[0] * w[1] - u[1] * w[0]};
      for (int d = 0; d < 3; ++d) sf[d] += cr[d] * 0.5;
    }
    double sn = 0;
    for (int d = 0; d < 3; ++d) {
      surf[d][f] = sf[d];
      sn += sf[d] * sf[d];
    }
    area[f] = std::sqrt(sn);
  };
  for (int kk = 0; kk < nz; ++kk)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        const long f = i + long(nx + 1) * (j + long(ny) * kk);
        const P3 nd[4] = {node(i, j, kk), node(i, j + 1, kk), node(i, j + 1, kk + 1), node(i, j, kk + 1)};
        face_geom(f, nd);
        fcm[f] = i > 0 ? cell(i - 1, j, kk) : -1;
        fcp[f] = i < nx ? cell(i, j, kk) : -1;
      }
  for (int kk = 0; kk < nz; ++kk)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long f = nfx + i + long(nx) * (j + long(ny + 1) * kk);
        const P3 nd[4] = {node(i, j, kk), node(i, j, kk + 1), node(i + 1, j, kk + 1), node(i + 1, j, kk)};
        face_geom(f, nd);
        fcm[f] = j > 0 ? cell(i, j - 1, kk) : -1;
        fcp[f] = j < ny ? cell(i, j, kk) : -1;
      }
  for (int kk = 0; kk <= nz; ++kk)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long f = nfx + nfy + i + long(nx) * (j + long(ny) * kk);
        const P3 nd[4] = {node(i, j, kk), node(i + 1, j, kk), node(i + 1, j + 1, kk), node(i, j + 1, kk)};
        face_geom(f, nd);
        fcm[f] = kk > 0 ? cell(i, j, kk - 1) : -1;
        fcp[f] = kk < nz ? cell(i, j, kk) : -1;
      }
  for (int kk = 0; kk < nz; ++kk)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long cl = cell(i, j, kk);
        const P3 nd[8] = {node(i, j, kk),         node(i + 1, j, kk),
                          node(i, j + 1, kk),     node(i + 1, j + 1, kk),
                          node(i, j, kk + 1),     node(i + 1, j, kk + 1),
                          node(i, j + 1, kk + 1), node(i + 1, j + 1, kk + 1)};
        double s3[3] = {0, 0, 0};
        for (int q = 0; q < 8; ++q)
          for (int d = 0; d < 3; ++d) s3[d] += nd[q].v[d];
        for (int d = 0; d < 3; ++d) cc[d][cl] = s3[d] / 8.;
      }
  auto face_of = [&](long cl, int q) {
    const int i = int(cl % nx), j = int((cl / nx) % ny), kk = int(cl / (long(nx) * ny));
    switch (q) {
      case 0: return i + long(nx + 1) * (j + long(ny) * kk);
      case 1: return i + 1 + long(nx + 1) * (j + long(ny) * kk);
      case 2: return nfx + i + long(nx) * (j + long(ny + 1) * kk);
      case 3: return nfx + i + long(nx) * (j + 1 + long(ny + 1) * kk);
      case 4: return nfx + nfy + i + long(nx) * (j + long(ny) * kk);
      default: return nfx + nfy + i + long(nx) * (j + long(ny) * (kk + 1));
    }
  };
  for (long cl = 0; cl < nc; ++cl) {
    double r = 0.;
    for (int q = 0; q < 6; ++q) {
      const long f = face_of(cl, q);
      r += fc[0][f] * (surf[0][f] * ((q % 2 == 0) ? -1. : 1.));
    }
    vol[cl] = r;
  }
  for (long f = 0; f < nf; ++f) {
    for (int d = 0; d < 3; ++d) {
      if (fcm[f] >= 0) t0[d][f] = cc[d][fcm[f]] - fc[d][f];
      if (fcp[f] >= 0) t1[d][f] = cc[d][fcp[f]] - fc[d][f];
      nrm[d][f] = surf[d][f] / area[f];
    }
  }
  auto up = [&](const std::vector<double>& h) {
    double* d = c->Alloc(long(h.size()));
    H3_CHECK(hipMemcpyAsync(d, h.data(), h.size() * 8, hipMemcpyHostToDevice, c->st));
    return (const double*)d;
  };
  G& g = c->g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.nc = nc;
  g.nfx = nfx;
  g.nfy = nfy;
  g.nf = nf;
  for (int d = 0; d < 3; ++d) {
    g.fc[d] = up(fc[d]);
    g.surf[d] = up(surf[d]);
    g.nrm[d] = up(nrm[d]);
    g.to0[d] = up(t0[d]);
    g.to1[d] = up(t1[d]);
    g.cc[d] = up(cc[d]);
  }
  g.area = up(area);
  g.vol = up(vol);
  c->Sync();  // the host vectors go out of scope
}

// Interpolate (the three boundary directions in face order)
void Interp(hg3_ctx* c, const double* u, double* res, const Bc& bc) {
  const G& g = c->g;
  // extrapolated boundary faces read the values of the cell's other faces
  // (earlier directions' boundary faces included): they keep the
  // direction-by-direction kernels; the other conditions are written with
  // the inner faces in one launch
  bool extrap = false;
  for (int s = 0; s < 6; ++s) extrap = extrap || bc.kind[s] == BC_EXTRAP;
  KInterpInner<<<Grid(g.nf), kB, 0, c->st>>>(g, u, res, bc, extrap ? 0 : 1);
  if (!extrap) return;
  const long per[3] = {2L * g.ny * g.nz, 2L * g.nx * g.nz, 2L * g.nx * g.ny};
  for (int d = 0; d < 3; ++d) KInterpBoundary<<<Grid(per[d]), kB, 0, c->st>>>(g, d, u, res, bc);
}
void Grad(hg3_ctx* c, const double* uf, double* const* out) {
  KGradient<<<Grid(c->nc), kB, 0, c->st>>>(c->g, uf, out[0], out[1], out[2]);
}
void Smooth(hg3_ctx* c, double* u, int repeat) {
  for (int r = 0; r < repeat; ++r) {
    Interp(c, u, c->ff, MakeBc(BC_DERIV, 0.));
    KAverage<<<Grid(c->nc), kB, 0, c->st>>>(c->g, c->ff, u);
  }
}
Bc VelBc(const hg3_ctx* c, int d) {
  Bc b;
  for (int s = 0; s < 6; ++s) {
    b.kind[s] = BC_VALUE;
    b.val[s] = c->cfg3.bc_velocity[s][d];
  }
  return b;
}
void Copy(hg3_ctx* c, double* dst, const double* src, long n) {
  H3_CHECK(hipMemcpyAsync(dst, src, size_t(n) * 8, hipMemcpyDeviceToDevice, c->st));
}

void UpdateProperties(hg3_ctx* c) {
  const hg_config& cfg = c->cfg;
  Phases3 ph{};
  ph.np = c->np;
  for (int i = 0; i < c->np; ++i) {
    ph.density[i] = cfg.density[i];
    ph.viscosity[i] = cfg.viscosity[i];
    ph.pd[i] = c->pd[i];
    ph.alpha[i] = c->alpha[i];
  }
  KProperties<<<Grid(c->nc), kB, 0, c->st>>>(c->nc, ph, c->rho, c->mu);
  Copy(c, c->rho_s, c->rho, c->nc);
  Smooth(c, c->rho_s, cfg.density_smooth_times);
  Copy(c, c->mu_s, c->mu, c->nc);
  Smooth(c, c->mu_s, cfg.viscosity_smooth_times);
  for (int d = 0; d < 3; ++d) {
    KForce<<<Grid(c->nc), kB, 0, c->st>>>(c->nc, c->rho, c->grav[d], c->force[d], c->f[d]);
    Smooth(c, c->f[d], cfg.force_smooth_times);
  }
}

void InitialFlux(hg3_ctx* c) {
  for (int d = 0; d < 3; ++d) Interp(c, c->u_tc[d], c->fu[d], VelBc(c, d));
  KFaceFlux<<<Grid(c->nf), kB, 0, c->st>>>(c->g, Faces3c{c->fu[0], c->fu[1], c->fu[2]}, c->fl_tc);
  Copy(c, c->fl_tp, c->fl_tc, c->nf);
}

// jacobi solve of rows R with constants k into x (returns iterations)
int SolveJacobi(hg3_ctx* c, const Rows& R, const double* k, double* x) {
  const hg_config& cfg = c->cfg;
  const int limit = cfg.lu_relaxed_num_iters_limit;
  const double tol = cfg.lu_relaxed_tolerance, omega = cfg.lu_relaxed_relaxation_factor;
  H3_CHECK(hipMemsetAsync(c->jdiff, 0, size_t(limit + 2) * 8, c->st));
  H3_CHECK(hipMemsetAsync(c->jstate, 0, 4 * sizeof(int), c->st));
  H3_CHECK(hipMemsetAsync(c->corr, 0, size_t(c->nc) * 8, c->st));
  double* buf[2] = {c->corr, c->corr2};
  int s = 0;
  while (true) {
    const int batch = std::min(32, limit + 1 - s);
    for (int b = 0; b < batch; ++b, ++s)
      KJacobiSweep<<<Grid(c->nc), kB, 0, c->st>>>(c->g, R, k, buf[s & 1], buf[(s + 1) & 1], omega,
                                                   c->jdiff, s, c->jstate, tol, limit);
    KJacobiCheck<<<1, 1, 0, c->st>>>(c->jdiff, s - 1, tol, limit, c->jstate);
    H3_CHECK(hipMemcpyAsync(c->state_h, c->jstate, 4 * sizeof(int), hipMemcpyDeviceToHost, c->st));
    c->Sync();
    if (c->state_h[0]) break;
    if (s > limit) throw Err3(HG_ERR_HIP, "jacobi: stop rule not reached");
  }
  Copy(c, x, buf[c->state_h[2]], c->nc);
  return c->state_h[1];
}

// The velocity components' jacobi solves together (KJacobiSweep3): each
// component's iterations and result are those of its own SolveJacobi; out
// u_ic[d] = u_ip[d] + x_d, iters[d] its sweep count.
void SolveVelocityJacobi3(hg3_ctx* c, const Rows& R, int iters[3]) {
  const hg_config& cfg = c->cfg;
  const int limit = cfg.lu_relaxed_num_iters_limit;
  const double tol = cfg.lu_relaxed_tolerance, omega = cfg.lu_relaxed_relaxation_factor;
  const int stride = limit + 2;
  H3_CHECK(hipMemsetAsync(c->jdiff, 0, size_t(3) * stride * 8, c->st));
  H3_CHECK(hipMemsetAsync(c->jstate, 0, 12 * sizeof(int), c->st));
  for (int d = 0; d < 3; ++d)
    H3_CHECK(hipMemsetAsync(c->jbuf[d][0], 0, size_t(c->nc) * 8, c->st));
  int s = 0;
  while (true) {
    const int batch = std::min(32, limit + 1 - s);
    for (int b = 0; b < batch; ++b, ++s) {
      Jac3 J;
      for (int d = 0; d < 3; ++d) {
        J.k[d] = c->kk[d];
        J.x[d] = c->jbuf[d][s & 1];
        J.xn[d] = c->jbuf[d][(s + 1) & 1];
      }
      KJacobiSweep3<<<Grid(c->nc), kB, 0, c->st>>>(c->g, R, J, omega, c->jdiff, stride, s,
                                                    c->jstate, tol, limit);
    }
    KJacobiCheck3<<<1, 64, 0, c->st>>>(c->jdiff, stride, s - 1, tol, limit, c->jstate);
    H3_CHECK(hipMemcpyAsync(c->state_h, c->jstate, 12 * sizeof(int), hipMemcpyDeviceToHost, c->st));
    c->Sync();
    if (c->state_h[0] && c->state_h[4] && c->state_h[8]) break;
    if (s > limit) throw Err3(HG_ERR_HIP, "jacobi: stop rule not reached");
  }
  for (int d = 0; d < 3; ++d) {
    iters[d] = c->state_h[4 * d + 1];
    KAddTo<<<Grid(c->nc), kB, 0, c->st>>>(c->nc, c->u_ip[d], c->jbuf[d][c->state_h[4 * d + 2]],
                                          c->u_ic[d]);
  }
}

int SolveCg(hg3_ctx* c, const Rows& R, const double* k, double* x) {
  const hg_config& cfg = c->cfg;
  const long n = c->nc;
  const int nb = int(std::min<long>(Grid(n), kCgBlocks));  // partial sums
  const unsigned ne = Grid(n);                              // elementwise
  const double tol = cfg.pressure_tolerance_gpu;
  const int max_iters = cfg.pressure_max_iters_gpu;
  CgScal* sc = c->cs;
  H3_CHECK(hipMemsetAsync(sc, 0, sizeof(CgScal), c->st));
  // b = -k - mean(-k)
  KSumK<<<nb, kB, 0, c->st>>>(n, k, c->part);
  KFold<<<1, kB, 0, c->st>>>(c->part, nb, &sc->mean);
  KDivide<<<1, 1, 0, c->st>>>(&sc->mean, double(n));
  KCgStart<<<nb, kB, 0, c->st>>>(n, k, &sc->mean, R.t[3], c->cb, c->cr, c->cz, c->cp, c->cx,
                                 c->part);
  KFold<<<1, kB, 0, c->st>>>(c->part, nb, &sc->rz);
  // ||b||^2
  KApplyRows<<<nb, kB, 0, c->st>>>(c->g, R, c->cx, c->cq, nullptr, nullptr);  // q = A 0 = 0
  KCgResidual<<<nb, kB, 0, c->st>>>(n, c->cb, c->cq, c->part);
  KFold<<<1, kB, 0, c->st>>>(c->part, nb, &sc->bb);
  H3_CHECK(hipMemcpyAsync(c->cs_h, sc, sizeof(CgScal), hipMemcpyDeviceToHost, c->st));
  c->Sync();
  int it = 0;
  if (c->cs_h->bb > 0.) {
    while (true) {
      for (int b = 0; b < 16; ++b) {
        KApplyRows<<<nb, kB, 0, c->st>>>(c->g, R, c->cp, c->cq, c->cp, c->part);
        KFold<<<1, kB, 0, c->st>>>(c->part, nb, &sc->pq);
        KCgUpdate<<<nb, kB, 0, c->st>>>(n, c->cp, c->cq, c->cx, c->cr, R.t[3], c->cz, sc,
                                        c->part, c->part + nb);
        KFold<<<1, kB, 0, c->st>>>(c->part, nb, &sc->rr);
        KFold<<<1, kB, 0, c->st>>>(c->part + nb, nb, &sc->rz2);
        KCgCheck<<<1, 1, 0, c->st>>>(sc, tol * tol, max_iters);
        KCgDir<<<ne, kB, 0, c->st>>>(n, c->cz, c->cp, sc);
        KCgShift<<<1, 1, 0, c->st>>>(sc);
      }
      H3_CHECK(hipMemcpyAsync(c->cs_h, sc, sizeof(CgScal), hipMemcpyDeviceToHost, c->st));
      c->Sync();
      if (c->cs_h->done) break;
    }
    it = c->cs_h->iters;
    c->p_converged = c->cs_h->converged != 0.;
  }
  Copy(c, x, c->cx, n);
  return it;
}

// fcg_mg: b = -k - mean(-k) as for cg, then FCG with the V-cycle; a
// breakdown of the preconditioner's descent property finishes with the
// Jacobi-PCG solve (from zero, to the same tolerance)
int SolveFcgMg(hg3_ctx* c, const Rows& R, const double* k, double* x) {
  const hg_config& cfg = c->cfg;
  const long n = c->nc;
  const int nb = int(std::min<long>(Grid(n), kCgBlocks));
  CgScal* sc = c->cs;
  KSumK<<<nb, kB, 0, c->st>>>(n, k, c->part);
  KFold<<<1, kB, 0, c->st>>>(c->part, nb, &sc->mean);
  KDivide<<<1, 1, 0, c->st>>>(&sc->mean, double(n));
  KFormB<<<Grid(n), kB, 0, c->st>>>(n, k, &sc->mean, c->cb);
  const hg3::Fcg3Result res = c->mg->Solve(R.t, c->cb, x, cfg.pressure_tolerance_gpu,
                                           cfg.pressure_max_iters_gpu, c->st);
  if (res.failed) return res.iterations + SolveCg(c, R, k, x);
  c->p_converged = res.converged;
  return res.iterations;
}

int RunSolver(hg3_ctx* c, int which, const Rows& R, const double* k, double* x) {
  if (which == HG_SOLVER_JACOBI) return SolveJacobi(c, R, k, x);
  if (which == HG_SOLVER_FCG_MG) return SolveFcgMg(c, R, k, x);
  return SolveCg(c, R, k, x);
}

void StartStep(hg3_ctx* c) {
  c->fluid_iter = 0;
  c->cd_iter = 0;
  H3_CHECK(hipMemsetAsync(c->flag, 0, sizeof(int), c->st));
  KNan<<<Grid(c->nc), kB, 0, c->st>>>(c->nc, c->p_tc, 1, c->flag);
  for (int d = 0; d < 3; ++d) KNan<<<Grid(c->nc), kB, 0, c->st>>>(c->nc, c->u_tc[d], 2, c->flag);
  H3_CHECK(hipMemcpyAsync(c->state_h, c->flag, sizeof(int), hipMemcpyDeviceToHost, c->st));
  c->Sync();
  if (c->state_h[0] & 1) throw Err3(HG_ERR_NAN, "NaN initial pressure");
  if (c->state_h[0] & 2) throw Err3(HG_ERR_NAN, "NaN initial field");
  const double ge = c->cfg.guess_extrapolation;
  for (int d = 0; d < 3; ++d)
    KExtrapolate<<<Grid(c->nc), kB, 0, c->st>>>(c->nc, c->u_tc[d], c->u_tp[d], ge, c->u_ic[d]);
  KExtrapolate<<<Grid(c->nc), kB, 0, c->st>>>(c->nc, c->p_tc, c->p_tp, ge, c->p_ic);
  KExtrapolate<<<Grid(c->nf), kB, 0, c->st>>>(c->nf, c->fl_tc, c->fl_tp, ge, c->fl_ic);
}

double Indicator(hg3_ctx* c) {
  const int nb = int(Grid(c->nc));
  KDiffMax<<<nb, kB, 0, c->st>>>(c->nc, c->u_ic[0], c->u_ic[1], c->u_ic[2], c->u_ip[0],
                                 c->u_ip[1], c->u_ip[2], c->part);
  KMaxFold<<<1, kB, 0, c->st>>>(c->part, nb, c->scal_d);
  H3_CHECK(hipMemcpyAsync(c->scal_h, c->scal_d, 8, hipMemcpyDeviceToHost, c->st));
  c->Sync();
  return c->scal_h[0];
}

void Iterate(hg3_ctx* c, hg_iter_stats* is) {
  const hg_config& cfg = c->cfg;
  const G& g = c->g;
  const long nc = c->nc, nf = c->nf;
  const Bc d0 = MakeBc(BC_DERIV, 0.);
  Copy(c, c->p_ip, c->p_ic, nc);
  Copy(c, c->fl_ip, c->fl_ic, nf);
  // restored force
  for (int d = 0; d < 3; ++d) Interp(c, c->f[d], c->ffe[d], d0);
  KRestoredForce<<<Grid(nc), kB, 0, c->st>>>(g, c->ffe[0], c->ffe[1], c->ffe[2], c->fr[0], c->fr[1],
                                             c->fr[2]);
  for (int d = 0; d < 3; ++d) Interp(c, c->fr[d], c->ffr[d], d0);
  Interp(c, c->mu_s, c->ffmu, d0);
  // pressure gradient
  Interp(c, c->p_ip, c->ffp, MakeBc(BC_EXTRAP));
  Grad(c, c->ffp, c->gp);
  for (int d = 0; d < 3; ++d) Interp(c, c->gp[d], c->ffgp[d], d0);
  // explicit viscosity
  for (int d = 0; d < 3; ++d) H3_CHECK(hipMemsetAsync(c->fo[d], 0, size_t(nc) * 8, c->st));
  for (int d = 0; d < 3; ++d) {
    Interp(c, c->u_ic[d], c->ff, VelBc(c, d));
    Grad(c, c->ff, c->g3);
    for (int e = 0; e < 3; ++e) Interp(c, c->g3[e], c->gf3[e], d0);
    KViscosity<<<Grid(nc), kB, 0, c->st>>>(g, d, c->ffmu, c->gf3[0], c->gf3[1], c->gf3[2],
                                           c->fo[0], c->fo[1], c->fo[2]);
  }
  KForceAssembly<<<Grid(nc), kB, 0, c->st>>>(
      g, c->rho_s, Force3{{c->gp[0], c->gp[1], c->gp[2]}, {c->fr[0], c->fr[1], c->fr[2]},
                          {c->u_ic[0], c->u_ic[1], c->u_ic[2]}, {c->fo[0], c->fo[1], c->fo[2]}});
  // momentum rows, component by component; coefficients from d = 0
  for (int d = 0; d < 3; ++d) {
    Copy(c, c->u_ip[d], c->u_ic[d], nc);
    Interp(c, c->u_ip[d], c->ff, VelBc(c, d));
    Grad(c, c->ff, c->g3);
    KMomentumFaces<<<Grid(nf), kB, 0, c->st>>>(g, c->fl_ip, c->ffmu, c->g3[0], c->g3[1], c->g3[2],
                                               VelBc(c, d), c->fe);
    MomCell m{c->rho_s, c->u_tp[d], c->u_tc[d], c->fo[d], c->u_ip[d], c->co[0], c->co[1], c->co[2],
              cfg.velocity_relaxation_factor};
    KMomentumRows<<<Grid(nc), kB, 0, c->st>>>(g, c->fe, m, d == 0 ? 1 : 0, c->A, c->kk[d]);
  }
  if (cfg.linear_solver_velocity == HG_SOLVER_JACOBI) {
    int it3[3];
    SolveVelocityJacobi3(c, c->A, it3);
    for (int d = 0; d < 3; ++d) is->velocity_solve_iters[d < 2 ? d : 1] = it3[d];
  } else {
    for (int d = 0; d < 3; ++d) {
      is->velocity_solve_iters[d < 2 ? d : 1] =
          RunSolver(c, cfg.linear_solver_velocity, c->A, c->kk[d], c->pc);
      KAddTo<<<Grid(nc), kB, 0, c->st>>>(nc, c->u_ip[d], c->pc, c->u_ic[d]);
    }
  }
  c->cd_iter += 1;
  // diagonal and Rhie-Chow
  KCoeffMean<<<Grid(nc), kB, 0, c->st>>>(g, c->A, c->a_p);
  Interp(c, c->a_p, c->ffd, MakeBc(BC_NONE));
  for (int d = 0; d < 3; ++d) Interp(c, c->u_ic[d], c->fu[d], VelBc(c, d));
  KRhieChow<<<Grid(nf), kB, 0, c->st>>>(g, Faces3{c->fu[0], c->fu[1], c->fu[2]},
                                        Faces3{c->ffgp[0], c->ffgp[1], c->ffgp[2]},
                                        Faces3{c->ffr[0], c->ffr[1], c->ffr[2]},
                                        Faces3{c->ffe[0], c->ffe[1], c->ffe[2]}, c->p_ip, c->ffd,
                                        cfg.rhie_chow_factor, c->fstar, c->fcoef);
  KPressureRows<<<Grid(nc), kB, 0, c->st>>>(g, c->fstar, c->fcoef, c->P);
  c->p_converged = true;  // jacobi: the reference's stop rule, no flag
  is->pressure_solve_iters = RunSolver(c, cfg.linear_solver_pressure, c->P, c->P.k, c->pc);
  is->pressure_converged = c->p_converged ? 1 : 0;
  // corrections
  KCorrectPressure<<<Grid(nc), kB, 0, c->st>>>(nc, c->p_ip, c->pc, cfg.pressure_relaxation_factor,
                                               c->p_ic);
  Interp(c, c->pc, c->ffpc, MakeBc(BC_EXTRAP));
  Grad(c, c->ffpc, c->g3);
  KCorrectVelocity<<<Grid(nc), kB, 0, c->st>>>(nc, c->g3[0], c->g3[1], c->g3[2], c->a_p, c->u_ic[0],
                                               c->u_ic[1], c->u_ic[2]);
  KCorrectFlux<<<Grid(nf), kB, 0, c->st>>>(g, c->fstar, c->fcoef, c->pc, c->fl_ic);
  c->fluid_iter += 1;
  is->iteration = c->fluid_iter;
  c->last_ind = Indicator(c);
  is->convergence_indicator = c->last_ind;
}

bool IsConverged(const hg3_ctx* c) {
  const double ind = c->cd_iter == 0 ? 1. : c->last_ind;
  return c->fluid_iter >= c->cfg.num_iterations_limit || ind < c->cfg.convergence_tolerance;
}

void FinishStep(hg3_ctx* c) {
  std::swap(c->p_tp, c->p_tc);
  std::swap(c->p_tc, c->p_ic);  // tp <- old tc, tc <- ic, ic <- old tp (free)
  std::swap(c->fl_tp, c->fl_tc);
  std::swap(c->fl_tc, c->fl_ic);
  for (int d = 0; d < 3; ++d) {
    std::swap(c->u_tp[d], c->u_tc[d]);
    std::swap(c->u_tc[d], c->u_ic[d]);
  }
  H3_CHECK(hipMemsetAsync(c->flag, 0, sizeof(int), c->st));
  KNan<<<Grid(c->nc), kB, 0, c->st>>>(c->nc, c->p_tc, 1, c->flag);
  for (int d = 0; d < 3; ++d) KNan<<<Grid(c->nc), kB, 0, c->st>>>(c->nc, c->u_tc[d], 2, c->flag);
  H3_CHECK(hipMemcpyAsync(c->state_h, c->flag, sizeof(int), hipMemcpyDeviceToHost, c->st));
  c->Sync();
  if (c->state_h[0] & 1) throw Err3(HG_ERR_NAN, "NaN pressure");
  if (c->state_h[0] & 2) throw Err3(HG_ERR_NAN, "NaN field");
  c->t_fluid += c->dt;
}

void AdvectSubstep(hg3_ctx* c) {
  const G& g = c->g;
  for (int i = 0; i < c->np; ++i) {
    double* u = c->pd[i];
    Interp(c, u, c->ff, MakeBc(BC_DERIV, 0.));
    Grad(c, c->ff, c->g3);
    // the flux of the step's last iteration: time_curr after FinishStep
    KAdvFaces<<<Grid(c->nf), kB, 0, c->st>>>(g, u, c->g3[0], c->g3[1], c->g3[2], c->fl_tc, c->fu[0]);
    KAdvUpdate<<<Grid(c->nc), kB, 0, c->st>>>(g, c->fu[0], c->fl_tc, c->dt_adv, u);
  }
  c->t_adv += c->dt_adv;
}

double* FieldPtr(hg3_ctx* c, int field, bool write, long* n) {
  *n = c->nc;
  if (field == HG_FIELD_VELOCITY_X) return c->u_tc[0];
  if (field == HG_FIELD_VELOCITY_Y) return c->u_tc[1];
  if (field == HG_FIELD_VELOCITY_Z) return c->u_tc[2];
  if (field == HG_FIELD_PRESSURE) return c->p_tc;
  if (field == HG_FIELD_VOLUME_FLUX) {
    *n = c->nf;
    return c->fl_tc;
  }
  if (field >= HG_FIELD_PARTIAL_DENSITY_0 && field < HG_FIELD_PARTIAL_DENSITY_0 + c->np)
    return c->pd[field - HG_FIELD_PARTIAL_DENSITY_0];
  if (write) return nullptr;
  if (field >= HG_FIELD_VOLUME_FRACTION_0 && field < HG_FIELD_VOLUME_FRACTION_0 + c->np)
    return c->alpha[field - HG_FIELD_VOLUME_FRACTION_0];
  if (field == HG_FIELD_DENSITY) return c->rho;
  if (field == HG_FIELD_DENSITY_SMOOTH) return c->rho_s;
  if (field == HG_FIELD_VISCOSITY_SMOOTH) return c->mu_s;
  if (field == HG_FIELD_FORCE_X) return c->f[0];
  if (field == HG_FIELD_FORCE_Y) return c->f[1];
  if (field == HG_FIELD_FORCE_Z) return c->f[2];
  return nullptr;
}

hg_status Guard3(hg3_ctx* c, const std::function<void()>& fn) {
  try {
    fn();
    return HG_OK;
  } catch (const Err3& e) {
    if (c) c->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    if (c) c->err = e.what();
    return HG_ERR_HIP;
  }
}

}  // namespace

extern "C" {

hg_status hg3_create(const hg3_config* cfg3, hg3_ctx** out) {
  if (!cfg3 || !out) return HG_ERR_BAD_ARG;
  *out = nullptr;
  const hg_config& cfg = cfg3->base;
  auto refuse = [&](const char* why) {
    g_create_error3 = why;
    return HG_ERR_UNSUPPORTED;
  };
  if (cfg.abi_version != HG_ABI_VERSION) return refuse("hg3_config from another ABI version");
  if (cfg.nx < 2 || cfg.ny < 2 || cfg3->nz < 2) return refuse("Nx, Ny, Nz must be >= 2");
  if (3.0 * (double(cfg.nx) + 1.) * (double(cfg.ny) + 1.) * (double(cfg3->nz) + 1.) >= 2147483648.)
    return refuse("the 3D path takes grids of fewer than 2^31 faces");
  if (cfg.num_phases < 1 || cfg.num_phases > 2) return refuse("num_phases must be 1 or 2");
  for (int s = 0; s < 6; ++s)
    if (cfg3->bc_kind[s] != HG_BC_WALL) return refuse("the 3D path runs walls only");
  if (cfg.heat_enable || cfg.simpler || cfg.dt_auto || cfg.tvd_split ||
      std::abs(cfg.sharp) > 1e-10 || (cfg.num_phases > 1 && cfg.sigma != 0.) ||
      cfg.meshvel[0] != 0. || cfg.meshvel[1] != 0. || cfg3->meshvel_z != 0. ||
      cfg.has_pressure_fixed_point || cfg.has_initial_sin)
    return refuse("heat, simpler, dt_auto, tvd_split, sharp, sigma, meshvel, a fixed pressure "
                  "point and initial_sin are not run on the 3D path");
  if (cfg.linear_solver_velocity != HG_SOLVER_JACOBI)
    return refuse("linear_solver_velocity: the 3D path runs jacobi");
  if (cfg.linear_solver_pressure != HG_SOLVER_JACOBI && cfg.linear_solver_pressure != HG_SOLVER_CG &&
      cfg.linear_solver_pressure != HG_SOLVER_FCG_MG)
    return refuse("linear_solver_pressure: the 3D path runs jacobi, cg and fcg_mg");
  if (cfg.linear_solver_pressure == HG_SOLVER_FCG_MG &&
      !hg3::Mg3Fcg::Supports(cfg.nx, cfg.ny, cfg3->nz))
    return refuse("linear_solver_pressure fcg_mg: Nx, Ny, Nz must halve (while even) to a "
                  "coarsest grid of at most 4096 cells");
  hg3_ctx* c = new hg3_ctx();
  const hg_status rc = Guard3(c, [&]() {
    c->cfg3 = *cfg3;
    c->cfg = cfg;
    c->np = cfg.num_phases;
    H3_CHECK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    BuildGeometry(c);
    const long nc = c->g.nc, nf = c->g.nf;
    c->nc = nc;
    c->nf = nf;
    for (int d = 0; d < 3; ++d) {
      c->u_tc[d] = c->Alloc(nc);
      c->u_tp[d] = c->Alloc(nc);
      c->u_ic[d] = c->Alloc(nc);
      c->u_ip[d] = c->Alloc(nc);
    }
    c->p_tc = c->Alloc(nc);
    c->p_tp = c->Alloc(nc);
    c->p_ic = c->Alloc(nc);
    c->p_ip = c->Alloc(nc);
    c->fl_tc = c->Alloc(nf);
    c->fl_tp = c->Alloc(nf);
    c->fl_ic = c->Alloc(nf);
    c->fl_ip = c->Alloc(nf);
    for (int i = 0; i < 2; ++i) {
      c->pd[i] = c->Alloc(nc);
      c->alpha[i] = c->Alloc(nc);
    }
    c->rho = c->Alloc(nc);
    c->rho_s = c->Alloc(nc);
    c->mu = c->Alloc(nc);
    c->mu_s = c->Alloc(nc);
    for (int d = 0; d < 3; ++d) {
      c->f[d] = c->Alloc(nc);
      c->ffe[d] = c->Alloc(nf);
      c->fr[d] = c->Alloc(nc);
      c->ffr[d] = c->Alloc(nf);
      c->gp[d] = c->Alloc(nc);
      c->ffgp[d] = c->Alloc(nf);
      c->fo[d] = c->Alloc(nc);
      c->g3[d] = c->Alloc(nc);
      c->gf3[d] = c->Alloc(nf);
      c->kk[d] = c->Alloc(nc);
      c->fu[d] = c->Alloc(nf);
    }
    c->ffmu = c->Alloc(nf);
    c->ffp = c->Alloc(nf);
    c->ff = c->Alloc(nf);
    c->fe = FaceExpr{c->Alloc(nf), c->Alloc(nf), c->Alloc(nf), c->Alloc(nf), c->Alloc(nf),
                     c->Alloc(nf)};
    for (int q = 0; q < 7; ++q) {
      c->A.t[q] = c->Alloc(nc);
      c->P.t[q] = c->Alloc(nc);
    }
    c->A.k = nullptr;
    c->P.k = c->Alloc(nc);
    c->corr = c->Alloc(nc);
    c->corr2 = c->Alloc(nc);
    c->a_p = c->Alloc(nc);
    c->ffd = c->Alloc(nf);
    c->fstar = c->Alloc(nf);
    c->fcoef = c->Alloc(nf);
    c->pc = c->Alloc(nc);
    c->ffpc = c->Alloc(nf);
    c->part = c->Alloc(2L * Grid(nc) + 2);
    c->scal_d = c->Alloc(4);
    c->cb = c->Alloc(nc);
    c->cr = c->Alloc(nc);
    c->cz = c->Alloc(nc);
    c->cp = c->Alloc(nc);
    c->cq = c->Alloc(nc);
    c->cx = c->Alloc(nc);
    c->cs = (CgScal*)c->Alloc(long(sizeof(CgScal) / 8 + 1));
    c->jlimit = std::max(1, cfg.lu_relaxed_num_iters_limit);
    // three components' max |corr| per sweep and states (done, iters, buffer)
    c->jdiff = (unsigned long long*)c->Alloc(3L * (c->jlimit + 2));
    c->jstate = (int*)c->Alloc(6);
    if (cfg.linear_solver_velocity == HG_SOLVER_JACOBI)
      for (int d = 0; d < 3; ++d)
        for (int b = 0; b < 2; ++b) c->jbuf[d][b] = c->Alloc(nc);
    c->flag = (int*)c->Alloc(1);
    H3_CHECK(hipHostMalloc(&c->state_h, 12 * sizeof(int)));  // three components' jacobi states
    H3_CHECK(hipHostMalloc(&c->scal_h, 4 * sizeof(double)));
    H3_CHECK(hipHostMalloc(&c->cs_h, sizeof(CgScal)));
    if (cfg.linear_solver_pressure == HG_SOLVER_FCG_MG)
      c->mg = std::make_unique<hg3::Mg3Fcg>(cfg.nx, cfg.ny, cfg3->nz);
    c->grav[0] = cfg.gravity[0];
    c->grav[1] = cfg.gravity[1];
    c->grav[2] = cfg3->gravity_z;
    c->force[0] = cfg.force[0];
    c->force[1] = cfg.force[1];
    c->force[2] = cfg3->force_z;
    c->dt = cfg.dt;
    c->dt_adv = cfg.dt * cfg.advection_dt_factor;
    {
      std::vector<double> args = {-2. * c->dt, -c->dt, 0.};
      std::vector<double> cc(3, 0.);
      const size_t skip = cfg.time_second_order ? 0 : 1;
      std::vector<double> cut(args.begin() + skip, args.end());
      std::vector<double> r = DerivCoeffs(0., cut);
      for (size_t i = 0; i < r.size(); ++i) cc[skip + i] = r[i];
      for (int i = 0; i < 3; ++i) c->co[i] = cc[i];
    }
    // initial fields (hydro2d.hpp:316-566)
    const double iv[3] = {cfg.initial_velocity[0], cfg.initial_velocity[1], cfg3->initial_velocity_z};
    std::vector<double> h(size_t(nc), 0.);
    for (int d = 0; d < 3; ++d) {
      std::fill(h.begin(), h.end(), iv[d]);
      H3_CHECK(hipMemcpyAsync(c->u_tc[d], h.data(), size_t(nc) * 8, hipMemcpyHostToDevice, c->st));
      Copy(c, c->u_tp[d], c->u_tc[d], nc);
    }
    c->Sync();
    InitialFlux(c);
    // partial densities: phase 1 in A1-B1 (cell centres, closed box),
    // smoothed; phase 0 the complement
    const size_t ncu = static_cast<size_t>(nc);
    std::vector<double> ccx(ncu), ccy(ncu), ccz(ncu);
    H3_CHECK(hipMemcpyAsync(ccx.data(), c->g.cc[0], size_t(nc) * 8, hipMemcpyDeviceToHost, c->st));
    H3_CHECK(hipMemcpyAsync(ccy.data(), c->g.cc[1], size_t(nc) * 8, hipMemcpyDeviceToHost, c->st));
    H3_CHECK(hipMemcpyAsync(ccz.data(), c->g.cc[2], size_t(nc) * 8, hipMemcpyDeviceToHost, c->st));
    c->Sync();
    std::vector<double> pd0(ncu), pd1(ncu, 0.);
    for (long cl = 0; cl < nc; ++cl) pd0[cl] = cfg.density[0] * cfg.initial_volume_fraction[0];
    if (c->np >= 2) {
      const double lo[3] = {cfg.a1[0], cfg.a1[1], cfg3->a1_z};
      const double hi[3] = {cfg.b1[0], cfg.b1[1], cfg3->b1_z};
      for (long cl = 0; cl < nc; ++cl) {
        pd1[cl] = cfg.density[1] * cfg.initial_volume_fraction[1];
        const double x[3] = {ccx[cl], ccy[cl], ccz[cl]};
        bool in = true;
        for (int d = 0; d < 3; ++d) in = in && !(x[d] < lo[d] || hi[d] < x[d]);
        if (in) pd1[cl] = cfg.density[1];
      }
      H3_CHECK(hipMemcpyAsync(c->pd[1], pd1.data(), size_t(nc) * 8, hipMemcpyHostToDevice, c->st));
      Smooth(c, c->pd[1], cfg.initial_volume_fraction_smooth_times);
      H3_CHECK(hipMemcpyAsync(pd1.data(), c->pd[1], size_t(nc) * 8, hipMemcpyDeviceToHost, c->st));
      c->Sync();
    }
    for (long cl = 0; cl < nc; ++cl) {
      double vs = 0.;
      if (c->np >= 2) vs += pd1[cl] / cfg.density[1];
      pd0[cl] = (1. - vs) * cfg.density[0];
    }
    H3_CHECK(hipMemcpyAsync(c->pd[0], pd0.data(), size_t(nc) * 8, hipMemcpyHostToDevice, c->st));
    UpdateProperties(c);
    c->Sync();
  });
  if (rc != HG_OK) {
    g_create_error3 = c->err;
    delete c;
    return rc;
  }
  *out = c;
  return HG_OK;
}

void hg3_destroy(hg3_ctx* c) { delete c; }

const char* hg3_last_error(const hg3_ctx* c) {
  return c ? c->err.c_str() : g_create_error3.c_str();
}

hg_status hg3_step(hg3_ctx* c, hg_step_stats* s) {
  if (!c) return HG_ERR_BAD_ARG;
  return Guard3(c, [&]() {
    hg_step_stats st{};
    StartStep(c);
    st.pressure_converged = 1;
    if (c->cfg.fluid_enable) {
      while (!IsConverged(c)) {
        hg_iter_stats is{};
        Iterate(c, &is);
        if (!is.pressure_converged) st.pressure_converged = 0;
        st.pressure_solve_iters_total += is.pressure_solve_iters;
        st.convergence_indicator = is.convergence_indicator;
      }
    }
    st.simple_iterations = c->fluid_iter;
    FinishStep(c);
    if (c->cfg.advection_enable) {
      while (c->t_adv < c->t_fluid - 0.5 * c->dt_adv) {
        AdvectSubstep(c);
        st.advection_substeps += 1;
      }
    }
    UpdateProperties(c);
    c->Sync();
    st.time = c->t_fluid;
    st.dt = c->dt;
    if (s) *s = st;
  });
}

hg_status hg3_get_field(hg3_ctx* c, int32_t field, double* out, size_t n) {
  if (!c || !out) return HG_ERR_BAD_ARG;
  long len = 0;
  double* d = FieldPtr(c, field, false, &len);
  if (!d || size_t(len) != n) return HG_ERR_BAD_ARG;
  return Guard3(c, [&]() {
    H3_CHECK(hipMemcpyAsync(out, d, n * 8, hipMemcpyDeviceToHost, c->st));
    c->Sync();
  });
}

hg_status hg3_set_field(hg3_ctx* c, int32_t field, const double* in, size_t n) {
  if (!c || !in) return HG_ERR_BAD_ARG;
  long len = 0;
  double* d = FieldPtr(c, field, true, &len);
  if (!d || size_t(len) != n) return HG_ERR_BAD_ARG;
  return Guard3(c, [&]() {
    H3_CHECK(hipMemcpyAsync(d, in, n * 8, hipMemcpyHostToDevice, c->st));
    double* prev = nullptr;
    if (field == HG_FIELD_VELOCITY_X) prev = c->u_tp[0];
    if (field == HG_FIELD_VELOCITY_Y) prev = c->u_tp[1];
    if (field == HG_FIELD_VELOCITY_Z) prev = c->u_tp[2];
    if (field == HG_FIELD_PRESSURE) prev = c->p_tp;
    if (field == HG_FIELD_VOLUME_FLUX) prev = c->fl_tp;
    if (prev) H3_CHECK(hipMemcpyAsync(prev, in, n * 8, hipMemcpyHostToDevice, c->st));
    c->Sync();
  });
}

}  // extern "C"
