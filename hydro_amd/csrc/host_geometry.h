// host_geometry.h -- uniform-mesh geometry for the kernels (host code).
//
// Follows InitUniformMesh (mesh.hpp:722-736) and MeshStructured's
// CalcCenter / CalcArea / CalcSurface / CalcVolume and vect-to-cell
// (mesh2d.hpp:215-245, 392-404).  Derives the per-direction / per-side
// scalars of hg::Geo from a representative cell and face, then checks, over
// the whole mesh, that every cell and face has bit-identical values; on
// power-of-two spacings they always do.  When they do not (spacings like
// 1/96) the kernels still use the representative values, which differ from
// the reference's per-cell geometry by at most a unit in the last place.
#pragma once

#include <cmath>
#include <vector>

#include "geometry.h"

namespace hg {

struct HostMesh {
  int nx = 0, ny = 0;
  std::vector<double> xn, yn;  // node coordinates
  bool uniform_exact = true;
  Geo geo{};

  static double Norm2(double x, double y) {
    double s = 0;
    s += x * x;
    s += y * y;
    return std::sqrt(s);
  }
  // Cell centre: mean of the four nodes (0,0),(1,0),(1,1),(0,1).
  void CellCenter(int i, int j, double& cx, double& cy) const {
    double sx = 0, sy = 0;
    sx += xn[i]; sy += yn[j];
    sx += xn[i + 1]; sy += yn[j];
    sx += xn[i + 1]; sy += yn[j + 1];
    sx += xn[i]; sy += yn[j + 1];
    cx = sx / 4.;
    cy = sy / 4.;
  }
  // x-face (i,j): nodes (i,j), (i,j+1).  y-face (i,j): nodes (i+1,j), (i,j).
  struct Face {
    double cx, cy, area, sx, sy;
  };
  Face XFaceGeo(int i, int j) const {
    return MakeFace(xn[i], yn[j], xn[i], yn[j + 1]);
  }
  Face YFaceGeo(int i, int j) const {
    return MakeFace(xn[i + 1], yn[j], xn[i], yn[j]);
  }
  static Face MakeFace(double x0, double y0, double x1, double y1) {
    Face f;
    double cx = 0, cy = 0;
    cx += x0; cy += y0;
    cx += x1; cy += y1;
    f.cx = cx / 2.;
    f.cy = cy / 2.;
    f.area = Norm2(x0 - x1, y0 - y1);
    double vx = x1 - x0, vy = y1 - y0;
    f.sx = vy;
    f.sy = -vx;
    return f;
  }
  double Volume(int i, int j) const {
    Face w = XFaceGeo(i, j), e = XFaceGeo(i + 1, j), s = YFaceGeo(i, j), n = YFaceGeo(i, j + 1);
    double r = 0.;
    r += w.cx * (w.sx * -1.);
    r += e.cx * (e.sx * 1.);
    r += s.cx * (s.sx * -1.);
    r += n.cx * (n.sx * 1.);
    return r;
  }

  // Derives Geo from cell (i,j) and its faces.
  Geo Derive(int i, int j) const {
    Geo g{};
    g.nx = nx;
    g.ny = ny;
    g.nfx = (nx + 1) * ny;
    g.V = Volume(i, j);
    Face w = XFaceGeo(i, j), s = YFaceGeo(i, j);
    g.sx = w.sx;
    g.sy = s.sy;
    g.ax = w.area;
    g.ay = s.area;
    double cx, cy, cxm, cym;
    CellCenter(i, j, cx, cy);
    CellCenter(i - 1 >= 0 ? i - 1 : i, j, cxm, cym);
    g.tx0 = cxm - w.cx;   // vect to cell 0 of the x-face between i-1 and i
    g.tx1 = cx - w.cx;
    double cxs, cys;
    CellCenter(i, j - 1 >= 0 ? j - 1 : j, cxs, cys);
    g.ty0 = cys - s.cy;
    g.ty1 = cy - s.cy;
    g.dnx = Norm2(g.tx1 - g.tx0, 0.);
    g.dny = Norm2(0., g.ty1 - g.ty0);
    g.rdist_x = Norm2(w.cx - cx, w.cy - cy);
    g.rdist_y = Norm2(s.cx - cx, s.cy - cy);
    return g;
  }

  void Init(int nx_, int ny_, const double a[2], const double b[2]) {
    nx = nx_;
    ny = ny_;
    xn.resize(nx + 1);
    yn.resize(ny + 1);
    double dx = b[0] - a[0], dy = b[1] - a[1];
    for (int i = 0; i <= nx; ++i) xn[i] = a[0] + (double(i) / double(nx)) * dx;
    for (int j = 0; j <= ny; ++j) yn[j] = a[1] + (double(j) / double(ny)) * dy;
    geo = Derive(1, 1);
    geo.invV = 1. / geo.V;
    int e = 0;
    geo.pow2V = std::frexp(geo.V, &e) == 0.5 ? 1 : 0;
    // Boundary sides.  The valid neighbour of a left/bottom face is cell 1
    // (factor -1), of a right/top face cell 0 (factor +1).
    for (int s = 0; s < 4; ++s) {
      int ci = (s == kRight) ? nx - 1 : 0;
      int cj = (s == kTop) ? ny - 1 : 0;
      double cx, cy;
      CellCenter(ci, cj, cx, cy);
      Face self, opp;
      if (s == kLeft) { self = XFaceGeo(0, cj); opp = XFaceGeo(1, cj); }
      else if (s == kRight) { self = XFaceGeo(nx, cj); opp = XFaceGeo(nx - 1, cj); }
      else if (s == kBottom) { self = YFaceGeo(ci, 0); opp = YFaceGeo(ci, 1); }
      else { self = YFaceGeo(ci, ny); opp = YFaceGeo(ci, ny - 1); }
      double factor = (s == kLeft || s == kBottom) ? -1. : 1.;
      double dist = Norm2(cx - self.cx, cy - self.cy);
      geo.bal[s] = 1. / dist * factor;
      geo.ex_dist[s] = dist;
      // outward normal of the domain at this face: normal * factor
      double nxv = (self.sx / self.area) * factor, nyv = (self.sy / self.area) * factor;
      // outward factors of the self / opposite faces seen from the cell
      double oself = (s == kLeft || s == kBottom) ? -1. : 1.;
      double oopp = -oself;
      double dself = 0.;
      dself += (self.sx * oself) * nxv;
      dself += (self.sy * oself) * nyv;
      double dopp = 0.;
      dopp += (opp.sx * oopp) * nxv;
      dopp += (opp.sy * oopp) * nyv;
      double V = Volume(ci, cj);
      geo.ex_dot[s] = dopp;
      geo.ex_den[s] = 1. / dist - dself / V;
    }
    // Exactness check over the mesh.
    uniform_exact = true;
    for (int j = 0; j < ny && uniform_exact; ++j)
      for (int i = 1; i < nx; ++i) {
        Face w = XFaceGeo(i, j);
        if (w.sx != geo.sx || w.area != geo.ax) { uniform_exact = false; break; }
      }
    for (int j = 1; j < ny && uniform_exact; ++j)
      for (int i = 0; i < nx; ++i) {
        Face s = YFaceGeo(i, j);
        if (s.sy != geo.sy || s.area != geo.ay) { uniform_exact = false; break; }
      }
    for (int j = 1; j < ny && uniform_exact; ++j)
      for (int i = 1; i < nx; ++i) {
        Geo g2 = Derive(i, j);
        if (g2.V != geo.V || g2.tx0 != geo.tx0 || g2.tx1 != geo.tx1 || g2.ty0 != geo.ty0 ||
            g2.ty1 != geo.ty1) {
          uniform_exact = false;
          break;
        }
      }
  }
};

}  // namespace hg
