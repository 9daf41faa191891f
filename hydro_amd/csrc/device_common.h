// device_common.h -- geometry and boundary helpers shared by the HIP kernels.
//
// The mesh is the reference's uniform structured 2D mesh (mesh.hpp:722-736,
// mesh2d.hpp:20-434).  All geometric quantities a kernel needs are constant
// per face direction / boundary side; the host derives them once with the
// reference's own formulas (host_geometry.h) and hands them to kernels by
// value in `Geo`.  Kernels written against these scalars evaluate every
// expression in the reference's operation order, dropping only additions of
// exact zeros (products with the zero component of an axis-aligned surface
// or normal), which cannot change a result except for the sign of a zero.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "geometry.h"

namespace hg {

// x / V, as a product when that is exact.
__device__ __forceinline__ double OverV(const Geo& g, double x) {
  return g.pow2V ? x * g.invV : x / g.V;
}

__device__ __forceinline__ int XFace(const Geo& g, int i, int j) { return i + (g.nx + 1) * j; }
__device__ __forceinline__ int YFace(const Geo& g, int i, int j) { return g.nfx + i + g.nx * j; }

// Inner-face interpolation u[cm]*(1-0.5) + u[cp]*0.5 (solver.hpp:415-427).
__device__ __forceinline__ double Mid(double um, double up) { return um * 0.5 + up * 0.5; }

// Extrapolation boundary condition (solver.hpp:446-464) on side `s`, with
// `uc` the cell value and `uopp` the value on the cell's opposite face.
__device__ __forceinline__ double Extrap(const Geo& g, int s, double uc, double uopp) {
  double nom = uc / g.ex_dist[s];
  nom += uopp * g.ex_dot[s] / g.V;
  return nom / g.ex_den[s];
}

// Face values of a cell field at the four faces (W, E, S, N) of cell (i,j).
// kind: 0 = value condition (bv per side), 1 = zero derivative,
// 2 = extrapolation.
// Index of the boundary face on side s at position q along it.
__device__ __forceinline__ int BoundaryIndex(const Geo& g, int s, int q) {
  return s == kLeft ? q : s == kRight ? g.ny + q : s == kBottom ? 2 * g.ny + q
                                                                : 2 * g.ny + g.nx + q;
}
// Signed distance factor of the zero-index side's valid cell: the valid
// neighbour of a left / bottom face is cell 1 (factor -1), of a right / top
// face cell 0 (factor +1).
__device__ __forceinline__ double SideFactor(int s) {
  return (s == kLeft || s == kBottom) ? -1. : 1.;
}

// The cell's value and its four neighbours' (W, E, S, N), loaded together;
// a missing neighbour reads the cell itself (its value is then not used).
struct Cross {
  double c, w, e, s, n;
};
__device__ __forceinline__ Cross LoadCross(const double* __restrict__ u, int nx, int ny, int i,
                                           int j) {
  const int c = i + nx * j;
  Cross x;
  x.c = u[c];
  x.w = u[i > 0 ? c - 1 : c];
  x.e = u[i < nx - 1 ? c + 1 : c];
  x.s = u[j > 0 ? c - nx : c];
  x.n = u[j < ny - 1 ? c + nx : c];
  return x;
}

// CellFaces (below) on values already loaded.
template <int kind>
__device__ __forceinline__ void CellFacesV(const Geo& g, const Cross& x, int i, int j,
                                           const double* bv, double f[4]) {
  const int nx = g.nx, ny = g.ny;
  const double uc = x.c;
  double fw = 0, fe = 0, fs = 0, fn = 0;
  if (i > 0) fw = Mid(x.w, uc);
  if (i < nx - 1) fe = Mid(uc, x.e);
  if (j > 0) fs = Mid(x.s, uc);
  if (j < ny - 1) fn = Mid(uc, x.n);
  if (kind == 0) {
    if (i == 0) fw = bv[kLeft];
    if (i == nx - 1) fe = bv[kRight];
    if (j == 0) fs = bv[kBottom];
    if (j == ny - 1) fn = bv[kTop];
  } else if (kind == 1) {
    if (i == 0) fw = uc;
    if (i == nx - 1) fe = uc;
    if (j == 0) fs = uc;
    if (j == ny - 1) fn = uc;
  } else {
    if (i == 0) fw = Extrap(g, kLeft, uc, fe);
    if (i == nx - 1) fe = Extrap(g, kRight, uc, fw);
    if (j == 0) fs = Extrap(g, kBottom, uc, fn);
    if (j == ny - 1) fn = Extrap(g, kTop, uc, fs);
  }
  f[0] = fw;
  f[1] = fe;
  f[2] = fs;
  f[3] = fn;
}
template <int kind>
__device__ __forceinline__ void CellFaces(const Geo& g, const double* __restrict__ u, int i,
                                          int j, const double* bv, double f[4]) {
  CellFacesV<kind>(g, LoadCross(u, g.nx, g.ny, i, j), i, j, bv, f);
}

// Gradient from the four face values (solver.hpp:658-677).
__device__ __forceinline__ void GradFromFaces(const Geo& g, const double f[4], double& gx,
                                              double& gy) {
  double sxv = (-g.sx) * f[0];
  sxv += g.sx * f[1];
  double syv = (-g.sy) * f[2];
  syv += g.sy * f[3];
  gx = OverV(g, sxv);
  gy = OverV(g, syv);
}

// Wave-level reductions (64 lanes).
__device__ __forceinline__ double WaveSum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double WaveMax(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double WaveMin(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

}  // namespace hg
