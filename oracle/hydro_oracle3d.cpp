// hydro_oracle3d.cpp -- CPU oracle of the hydro3d time step (TEST
// INFRASTRUCTURE).
//
// The 3D restatement of the reference algorithm that hydro_oracle.cpp states
// in 2D: the reference's hydro<Mesh> is written once for both meshes
// (hydro2d.hpp, fluid.hpp, conv_diff.hpp, advection.hpp, solver.hpp,
// linear.hpp are templates over Mesh) and hydro3d.cpp:17-20 instantiates it
// on geom3d::MeshStructured (mesh3d.hpp:20-464).  Each function here follows
// its 2D counterpart with six faces per cell (W, E, S, N, B, F: the order of
// GetNeighbourFace, mesh3d.hpp:306-312) and 7-point rows whose terms are
// kept in cell-index order (B, S, W, C, E, N, F), as Expression sorts them.
// The mesh geometry is mesh3d.hpp's: face centres as the mean of the four
// nodes, surfaces as the sum of two triangle cross products, volumes from
// face centres and outward surfaces.
//
// Scope: one or two phases, walls on every side, gravity and force; the
// solvers lu, gauss_seidel, jacobi and cg.  Only tests/ load this library.

#include "hydro_oracle3d.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

using Vec = std::vector<double>;

struct V3 {
  double x, y, z;
  double operator[](int d) const { return d == 0 ? x : d == 1 ? y : z; }
};

// Vect::dot, norm (common/vect.hpp:173-189): sums from 0 in component order.
inline double Dot(V3 a, V3 b) {
  double s = 0;
  s += a.x * b.x;
  s += a.y * b.y;
  s += a.z * b.z;
  return s;
}
inline double Norm(V3 a) {
  double s = 0;
  s += a.x * a.x;
  s += a.y * a.y;
  s += a.z * a.z;
  return std::sqrt(s);
}
inline V3 Sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
// Vect::cross (common/vect.hpp:193-199)
inline V3 Cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

enum BcKind { BC_NONE = 0, BC_VALUE, BC_DERIV, BC_EXTRAP };

struct SideBc {
  BcKind kind[6];
  double val[6];
};

SideBc MakeBc(BcKind k, double v = 0.) {
  SideBc b;
  for (int s = 0; s < 6; ++s) {
    b.kind[s] = k;
    b.val[s] = v;
  }
  return b;
}

// Structured 3D mesh (mesh.hpp:722-736 InitUniformMesh, mesh3d.hpp:240-460).
struct Mesh {
  int nx = 0, ny = 0, nz = 0;
  size_t nc = 0, nfx = 0, nfy = 0, nf = 0;
  std::vector<V3> cc;
  Vec vol;
  std::vector<V3> fc, surf, normal, to0, to1;
  Vec area;
  std::vector<long> fcm, fcp;
  std::vector<int> side;  // boundary side of a face (-1: inner)

  size_t Cell(int i, int j, int k) const { return size_t(i) + size_t(nx) * (j + size_t(ny) * k); }
  size_t FX(int i, int j, int k) const { return size_t(i) + size_t(nx + 1) * (j + size_t(ny) * k); }
  size_t FY(int i, int j, int k) const {
    return nfx + size_t(i) + size_t(nx) * (j + size_t(ny + 1) * k);
  }
  size_t FZ(int i, int j, int k) const {
    return nfx + nfy + size_t(i) + size_t(nx) * (j + size_t(ny) * k);
  }
  // neighbour face q of cell c: W, E, S, N, B, F (mesh3d.hpp:306-312)
  size_t Face(size_t c, int q) const {
    const int i = int(c % nx), j = int((c / nx) % ny), k = int(c / (size_t(nx) * ny));
    switch (q) {
      case 0: return FX(i, j, k);
      case 1: return FX(i + 1, j, k);
      case 2: return FY(i, j, k);
      case 3: return FY(i, j + 1, k);
      case 4: return FZ(i, j, k);
      default: return FZ(i, j, k + 1);
    }
  }
  // GetOutwardFactor (mesh3d.hpp:120-131)
  static double Out(int q) { return (q % 2 == 0) ? -1. : 1.; }
  V3 OutSurf(size_t c, int q) const {
    V3 s = surf[Face(c, q)];
    double f = Out(q);
    return V3{s.x * f, s.y * f, s.z * f};
  }
  bool Inner(size_t f) const { return fcm[f] >= 0 && fcp[f] >= 0; }
  int ValidId(size_t f) const { return fcm[f] >= 0 ? 0 : 1; }
  V3 ToCell(size_t f, int id) const { return id == 0 ? to0[f] : to1[f]; }

  void Init(int nx_, int ny_, int nz_, const double a[3], const double b[3]) {
    nx = nx_;
    ny = ny_;
    nz = nz_;
    nc = size_t(nx) * ny * nz;
    nfx = size_t(nx + 1) * ny * nz;
    nfy = size_t(nx) * (ny + 1) * nz;
    nf = nfx + nfy + size_t(nx) * ny * (nz + 1);
    const int nn[3] = {nx, ny, nz};
    // nodes: lb + (midx / size) * (rt - lb), component-wise
    auto node = [&](int i, int j, int k) {
      const int m[3] = {i, j, k};
      double r[3];
      for (int d = 0; d < 3; ++d) r[d] = a[d] + (double(m[d]) / double(nn[d])) * (b[d] - a[d]);
      return V3{r[0], r[1], r[2]};
    };
    fc.resize(nf);
    surf.resize(nf);
    normal.resize(nf);
    area.resize(nf);
    to0.assign(nf, V3{0, 0, 0});
    to1.assign(nf, V3{0, 0, 0});
    fcm.resize(nf);
    fcp.resize(nf);
    side.assign(nf, -1);
    // CalcCenter (mean of the four nodes), CalcSurface (triangles 0-1-2 and
    // 0-2-3, (b - a) x (c - a) * 0.5 each), CalcArea = |surface|
    auto face_geom = [&](size_t f, const V3 (&nd)[4]) {
      V3 c{0, 0, 0};
      for (int q = 0; q < 4; ++q) {
        c.x += nd[q].x;
        c.y += nd[q].y;
        c.z += nd[q].z;
      }
      fc[f] = V3{c.x / 4., c.y / 4., c.z / 4.};
      V3 s{0, 0, 0};
      for (int q = 2; q < 4; ++q) {
        V3 t = Cross(Sub(nd[q - 1], nd[0]), Sub(nd[q], nd[0]));
        s.x += t.x * 0.5;
        s.y += t.y * 0.5;
        s.z += t.z * 0.5;
      }
      surf[f] = s;
      area[f] = Norm(s);
    };
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {
          const size_t f = FX(i, j, k);
          const V3 nd[4] = {node(i, j, k), node(i, j + 1, k), node(i, j + 1, k + 1),
                            node(i, j, k + 1)};
          face_geom(f, nd);
          fcm[f] = i > 0 ? long(Cell(i - 1, j, k)) : -1;
          fcp[f] = i < nx ? long(Cell(i, j, k)) : -1;
          if (i == 0) side[f] = 0;
          if (i == nx) side[f] = 1;
        }
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const size_t f = FY(i, j, k);
          const V3 nd[4] = {node(i, j, k), node(i, j, k + 1), node(i + 1, j, k + 1),
                            node(i + 1, j, k)};
          face_geom(f, nd);
          fcm[f] = j > 0 ? long(Cell(i, j - 1, k)) : -1;
          fcp[f] = j < ny ? long(Cell(i, j, k)) : -1;
          if (j == 0) side[f] = 2;
          if (j == ny) side[f] = 3;
        }
    for (int k = 0; k <= nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const size_t f = FZ(i, j, k);
          const V3 nd[4] = {node(i, j, k), node(i + 1, j, k), node(i + 1, j + 1, k),
                            node(i, j + 1, k)};
          face_geom(f, nd);
          fcm[f] = k > 0 ? long(Cell(i, j, k - 1)) : -1;
          fcp[f] = k < nz ? long(Cell(i, j, k)) : -1;
          if (k == 0) side[f] = 4;
          if (k == nz) side[f] = 5;
        }
    cc.resize(nc);
    vol.resize(nc);
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const size_t c = Cell(i, j, k);
          // the eight nodes in cell_neighbour_node_offset_ order
          const V3 nd[8] = {node(i, j, k),         node(i + 1, j, k),
                            node(i, j + 1, k),     node(i + 1, j + 1, k),
                            node(i, j, k + 1),     node(i + 1, j, k + 1),
                            node(i, j + 1, k + 1), node(i + 1, j + 1, k + 1)};
          V3 s{0, 0, 0};
          for (int q = 0; q < 8; ++q) {
            s.x += nd[q].x;
            s.y += nd[q].y;
            s.z += nd[q].z;
          }
          cc[c] = V3{s.x / 8., s.y / 8., s.z / 8.};
        }
    for (size_t c = 0; c < nc; ++c) {
      double r = 0.;
      for (int q = 0; q < 6; ++q) r += fc[Face(c, q)].x * OutSurf(c, q).x;
      vol[c] = r;
    }
    for (size_t f = 0; f < nf; ++f) {
      if (fcm[f] >= 0) to0[f] = Sub(cc[fcm[f]], fc[f]);
      if (fcp[f] >= 0) to1[f] = Sub(cc[fcp[f]], fc[f]);
      normal[f] = V3{surf[f].x / area[f], surf[f].y / area[f], surf[f].z / area[f]};
    }
  }
};

// Interpolate (solver.hpp:392-470), one scalar field; boundary faces in face
// order, the extrapolation reading the values other faces hold by then.
void Interpolate(const Mesh& m, const double* u, double* res, const SideBc& bc) {
  for (size_t f = 0; f < m.nf; ++f)
    res[f] = m.Inner(f) ? u[m.fcm[f]] * (1. - 0.5) + u[m.fcp[f]] * 0.5 : 0.;
  for (size_t f = 0; f < m.nf; ++f) {
    if (m.Inner(f)) continue;
    const int side = m.side[f];
    const BcKind kind = bc.kind[side];
    if (kind == BC_NONE) continue;
    const int id = m.ValidId(f);
    const size_t c = size_t(id == 0 ? m.fcm[f] : m.fcp[f]);
    const double factor = (id == 0 ? 1. : -1.);
    if (kind == BC_VALUE) {
      res[f] = bc.val[side];
    } else if (kind == BC_DERIV) {
      const double alpha = Norm(m.ToCell(f, id)) * factor;
      res[f] = u[c] + bc.val[side] * alpha;
    } else {
      const V3 nrm{m.normal[f].x * factor, m.normal[f].y * factor, m.normal[f].z * factor};
      const double dist = Norm(m.ToCell(f, id));
      double nom = u[c] / dist;
      double den = 1. / dist;
      const double volume = m.vol[c];
      for (int q = 0; q < 6; ++q) {
        const size_t nface = m.Face(c, q);
        if (nface == f)
          den -= Dot(m.OutSurf(c, q), nrm) / volume;
        else
          nom += res[nface] * Dot(m.OutSurf(c, q), nrm) / volume;
      }
      res[f] = nom / den;
    }
  }
}

// Gradient (solver.hpp:658-677)
void Gradient(const Mesh& m, const double* uf, double* g[3]) {
  for (size_t c = 0; c < m.nc; ++c) {
    V3 s{0, 0, 0};
    for (int q = 0; q < 6; ++q) {
      const V3 o = m.OutSurf(c, q);
      const double u = uf[m.Face(c, q)];
      s.x += o.x * u;
      s.y += o.y * u;
      s.z += o.z * u;
    }
    g[0][c] = s.x / m.vol[c];
    g[1][c] = s.y / m.vol[c];
    g[2][c] = s.z / m.vol[c];
  }
}

// GetSmoothField (solver.hpp:636-656), Average over the six faces
void Smooth(const Mesh& m, Vec& u, int repeat) {
  const SideBc bc = MakeBc(BC_DERIV, 0.);
  Vec uf(m.nf);
  for (int r = 0; r < repeat; ++r) {
    Interpolate(m, u.data(), uf.data(), bc);
    for (size_t c = 0; c < m.nc; ++c) {
      double s = 0;
      for (int q = 0; q < 6; ++q) s += uf[m.Face(c, q)];
      u[c] = s / 6.;
    }
  }
}

double Superbee(double p, double q) {
  if (p > 0. && q > 0.) return std::max(std::min(2 * p, q), std::min(p, 2 * q));
  if (p < 0. && q < 0.) return -std::max(std::min(-2 * p, -q), std::min(-p, -2 * q));
  return 0.;
}

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
std::vector<double> DerivCoeffs(double target, const std::vector<double>& args, size_t skip) {
  std::vector<double> cut(args.begin() + skip, args.end());
  std::vector<double> cc = DerivCoeffs(target, cut);
  std::vector<double> res(args.size(), 0.);
  for (size_t i = 0; i < cc.size(); ++i) res[skip + i] = cc[i];
  return res;
}

// ---- 7-point rows: terms in cell-index order B, S, W, C, E, N, F
// (offsets -nx ny, -nx, -1, 0, +1, +nx, +nx ny)
struct Sys {
  int nx = 0, ny = 0, nz = 0;
  Vec t[7];  // coefficients, index order
  Vec k;
  void Resize(int nx_, int ny_, int nz_) {
    nx = nx_;
    ny = ny_;
    nz = nz_;
    for (auto& v : t) v.assign(size(), 0.);
    k.assign(size(), 0.);
  }
  size_t size() const { return size_t(nx) * ny * nz; }
  long Off(int q) const {
    const long nxy = long(nx) * ny;
    const long o[7] = {-nxy, -long(nx), -1, 0, 1, long(nx), nxy};
    return o[q];
  }
  bool Has(size_t r, int q) const {
    const int i = int(r % nx), j = int((r / nx) % ny), kk = int(r / (size_t(nx) * ny));
    switch (q) {
      case 0: return kk > 0;
      case 1: return j > 0;
      case 2: return i > 0;
      case 3: return true;
      case 4: return i < nx - 1;
      case 5: return j < ny - 1;
      default: return kk < nz - 1;
    }
  }
  // Expression::Evaluate (linear.hpp:128-135)
  double Eval(size_t r, const double* x, const double* kk) const {
    double v = kk[r];
    for (int q = 0; q < 7; ++q)
      if (Has(r, q)) v += x[long(r) + Off(q)] * t[q][r];
    return v;
  }
  // Expression::CoeffSum (linear.hpp:137-142)
  double CoeffSum(size_t r) const {
    double v = 0.;
    for (int q = 0; q < 7; ++q)
      if (Has(r, q)) v += t[q][r];
    return v;
  }
};
// face q (W, E, S, N, B, F) -> row term (index order)
constexpr int kFaceTerm[6] = {2, 4, 1, 5, 0, 6};

struct SolveResult {
  int iters = 0;
  double diff = 0.;
  bool converged = true;
};

// LuDecomposition::Solve (linear.hpp:527-567): forward over the terms below
// the diagonal in index order, backward over those above in reverse order
void SolveLu(const Sys& a, const Vec& k, Vec& x) {
  const size_t N = a.size();
  x.assign(N, 0.);
  for (size_t i = 0; i < N; ++i) {
    double sum = 0;
    for (int q = 0; q < 3; ++q)
      if (a.Has(i, q)) sum += a.t[q][i] * x[long(i) + a.Off(q)];
    x[i] = (-k[i] - sum) / a.t[3][i];
  }
  for (size_t i = N; i > 0;) {
    --i;
    double sum = 0;
    for (int q = 6; q > 3; --q)
      if (a.Has(i, q)) sum += a.t[q][i] * x[long(i) + a.Off(q)];
    x[i] -= sum / a.t[3][i];
  }
}

// GaussSeidel::Solve (linear.hpp:685-715) and Jacobi::Solve (:750-783):
// the off-diagonal terms summed in index order
template <bool GS>
SolveResult SolveRelax(const Sys& a, const Vec& k, Vec& x, double tol, int limit, double omega) {
  const size_t N = a.size();
  x.assign(N, 0.);
  Vec next = x;
  size_t iter = 0;
  double diff = 0.;
  do {
    diff = 0.;
    for (size_t i = 0; i < N; ++i) {
      double sum = 0.;
      for (int q = 0; q < 7; ++q)
        if (q != 3 && a.Has(i, q)) sum += a.t[q][i] * x[long(i) + a.Off(q)];
      const double value = -(k[i] + sum) / a.t[3][i];
      const double corr = value - x[i];
      diff = std::max(diff, std::abs(corr));
      if (GS)
        x[i] += corr * omega;
      else
        next[i] = x[i] + corr * omega;
    }
    if (!GS) std::swap(x, next);
  } while (diff > tol && iter++ < size_t(limit));
  SolveResult r;
  r.iters = int(iter);
  r.diff = diff;
  r.converged = !(diff > tol);
  return r;
}

// Jacobi-PCG to a relative residual (the GPU's cg, FP64), b = -k projected
// to zero mean (the pressure system's constants are its null space)
SolveResult SolveCg(const Sys& a, const Vec& k, Vec& x, double tol, int max_iters) {
  const size_t N = a.size();
  auto apply = [&](const Vec& v, Vec& y) {
    for (size_t r = 0; r < N; ++r) {
      double s = 0.;
      for (int q = 0; q < 7; ++q)
        if (a.Has(r, q)) s += a.t[q][r] * v[long(r) + a.Off(q)];
      y[r] = s;
    }
  };
  auto dot = [&](const Vec& u, const Vec& v) {
    double s = 0.;
    for (size_t r = 0; r < N; ++r) s += u[r] * v[r];
    return s;
  };
  Vec b(N);
  for (size_t i = 0; i < N; ++i) b[i] = -k[i];
  double mean = 0.;
  for (double v : b) mean += v;
  mean /= double(N);
  for (double& v : b) v -= mean;
  x.assign(N, 0.);
  SolveResult res;
  const double bnorm = std::sqrt(dot(b, b));
  if (bnorm == 0.) return res;
  Vec r = b, z(N), p(N), q(N);
  for (size_t i = 0; i < N; ++i) z[i] = r[i] / a.t[3][i];
  p = z;
  double rz = dot(r, z);
  int it = 0;
  while (it < max_iters) {
    apply(p, q);
    const double alpha = rz / dot(p, q);
    for (size_t i = 0; i < N; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    ++it;
    if (std::sqrt(dot(r, r)) / bnorm <= tol) break;
    for (size_t i = 0; i < N; ++i) z[i] = r[i] / a.t[3][i];
    const double rz2 = dot(r, z);
    const double beta = rz2 / rz;
    rz = rz2;
    for (size_t i = 0; i < N; ++i) p[i] = z[i] + beta * p[i];
  }
  apply(x, q);
  double t = 0.;
  for (size_t i = 0; i < N; ++i) t += (b[i] - q[i]) * (b[i] - q[i]);
  res.iters = it;
  res.diff = std::sqrt(t) / bnorm;
  res.converged = res.diff <= tol;
  return res;
}

}  // namespace

struct or3_ctx {
  hg3_config cfg3;
  hg_config cfg;  // cfg3.base
  Mesh m;
  int np = 1;
  double dt = 0., dt_adv = 0., t_fluid = 0., t_adv = 0.;
  std::vector<double> bdf;
  Vec u_tc[3], u_tp[3], u_ic[3], u_ip[3];
  Vec p_tc, p_tp, p_ic, p_ip;
  Vec fl_tc, fl_tp, fl_ic, fl_ip;
  Vec rho, rho_s, mu_s, f[3];
  Vec pd_tc[2], pd_tp[2], pd_ic[2];
  Vec alpha[2];
  int fluid_iter = 0, cd_iter = 0;
  double g[3] = {0, 0, 0}, force[3] = {0, 0, 0};
};

namespace {

SideBc VelBc(const or3_ctx* o, int comp) {
  SideBc b;
  for (int s = 0; s < 6; ++s) {
    b.kind[s] = BC_VALUE;
    b.val[s] = o->cfg3.bc_velocity[s][comp];
  }
  return b;
}

void InitialFlux(or3_ctx* o) {
  const Mesh& m = o->m;
  Vec uf[3];
  for (int d = 0; d < 3; ++d) {
    uf[d].resize(m.nf);
    Interpolate(m, o->u_tc[d].data(), uf[d].data(), VelBc(o, d));
  }
  o->fl_tc.assign(m.nf, 0.);
  const V3 mv{0., 0., 0.};
  for (size_t f = 0; f < m.nf; ++f) o->fl_tc[f] = Dot(V3{uf[0][f], uf[1][f], uf[2][f]}, m.surf[f]);
  for (size_t f = 0; f < m.nf; ++f) o->fl_tc[f] -= Dot(mv, m.surf[f]);
  o->fl_tp = o->fl_tc;
}

// hydro::UpdateFluidProperties (hydro2d.hpp:1404-1430), no sources
void UpdateProperties(or3_ctx* o) {
  const Mesh& m = o->m;
  const hg_config& cfg = o->cfg;
  const size_t nc = m.nc;
  for (size_t c = 0; c < nc; ++c) {
    double sum = 0.;
    for (int i = 0; i < o->np; ++i) {
      o->alpha[i][c] = o->pd_tc[i][c] / cfg.density[i];
      sum += o->alpha[i][c];
    }
    for (int i = 0; i < o->np; ++i) o->alpha[i][c] /= sum;
  }
  o->rho.assign(nc, 0.);
  Vec mu(nc, 0.);
  for (int i = 0; i < o->np; ++i)
    for (size_t c = 0; c < nc; ++c) o->rho[c] += cfg.density[i] * o->alpha[i][c];
  for (int i = 0; i < o->np; ++i)
    for (size_t c = 0; c < nc; ++c) mu[c] += cfg.viscosity[i] * o->alpha[i][c];
  o->rho_s = o->rho;
  Smooth(m, o->rho_s, cfg.density_smooth_times);
  o->mu_s = mu;
  Smooth(m, o->mu_s, cfg.viscosity_smooth_times);
  // CalcForce (hydro2d.hpp:1308-1376): gravity * rho + force
  for (int d = 0; d < 3; ++d) {
    o->f[d].resize(nc);
    for (size_t c = 0; c < nc; ++c) o->f[d][c] = o->g[d] * o->rho[c] + o->force[d];
  }
  for (int d = 0; d < 3; ++d) Smooth(m, o->f[d], cfg.force_smooth_times);
}

SolveResult RunSolver(const or3_ctx* o, int which, const Sys& a, const Vec& k, Vec& x) {
  const hg_config& cfg = o->cfg;
  SolveResult r;
  switch (which) {
    case HG_SOLVER_LU: SolveLu(a, k, x); break;
    case HG_SOLVER_GAUSS_SEIDEL:
      r = SolveRelax<true>(a, k, x, cfg.lu_relaxed_tolerance, cfg.lu_relaxed_num_iters_limit,
                           cfg.lu_relaxed_relaxation_factor);
      break;
    case HG_SOLVER_JACOBI:
      r = SolveRelax<false>(a, k, x, cfg.lu_relaxed_tolerance, cfg.lu_relaxed_num_iters_limit,
                            cfg.lu_relaxed_relaxation_factor);
      break;
    default:
      r = SolveCg(a, k, x, cfg.pressure_tolerance_gpu, cfg.pressure_max_iters_gpu);
      break;
  }
  return r;
}

// FluidSimple::StartStep (fluid.hpp:793-812)
int StartStep(or3_ctx* o) {
  const double ge = o->cfg.guess_extrapolation;
  o->fluid_iter = 0;
  o->cd_iter = 0;
  for (double v : o->p_tc)
    if (!(v * 0. == 0.)) return HG_ERR_NAN;
  for (int d = 0; d < 3; ++d) {
    for (double v : o->u_tc[d])
      if (!(v * 0. == 0.)) return HG_ERR_NAN;
    o->u_ic[d] = o->u_tc[d];
    for (size_t c = 0; c < o->m.nc; ++c) o->u_ic[d][c] += (o->u_tc[d][c] - o->u_tp[d][c]) * ge;
  }
  o->p_ic = o->p_tc;
  o->fl_ic = o->fl_tc;
  for (size_t c = 0; c < o->m.nc; ++c) o->p_ic[c] += (o->p_tc[c] - o->p_tp[c]) * ge;
  for (size_t f = 0; f < o->m.nf; ++f) o->fl_ic[f] += (o->fl_tc[f] - o->fl_tp[f]) * ge;
  return HG_OK;
}

double ConvergenceIndicator(const or3_ctx* o) {
  if (o->cd_iter == 0) return 1.;
  double r = 0.;
  for (size_t c = 0; c < o->m.nc; ++c) {
    const V3 a{o->u_ic[0][c], o->u_ic[1][c], o->u_ic[2][c]};
    const V3 b{o->u_ip[0][c], o->u_ip[1][c], o->u_ip[2][c]};
    r = std::max(r, Norm(Sub(b, a)));
  }
  return r;
}

// FluidSimple::MakeIteration (fluid.hpp:814-1158)
int Iterate(or3_ctx* o, hg_iter_stats* stats) {
  const hg_config& cfg = o->cfg;
  const Mesh& m = o->m;
  const size_t nc = m.nc, nf = m.nf;
  o->p_ip = o->p_ic;
  o->fl_ip = o->fl_ic;
  const SideBc d0 = MakeBc(BC_DERIV, 0.);

  // CalcExtForce (fluid.hpp:602-631): restored force from the face force
  Vec ffe[3], fr[3], ffr[3];
  for (int d = 0; d < 3; ++d) {
    ffe[d].resize(nf);
    Interpolate(m, o->f[d].data(), ffe[d].data(), d0);
  }
  for (int d = 0; d < 3; ++d) fr[d].resize(nc);
  for (size_t c = 0; c < nc; ++c) {
    V3 sum{0, 0, 0};
    for (int q = 0; q < 6; ++q) {
      const size_t fq = m.Face(c, q);
      const double dn = Dot(V3{ffe[0][fq], ffe[1][fq], ffe[2][fq]}, m.normal[fq]);
      const double dist = Norm(Sub(m.fc[fq], m.cc[c]));
      sum.x += m.surf[fq].x * dn * dist;
      sum.y += m.surf[fq].y * dn * dist;
      sum.z += m.surf[fq].z * dn * dist;
    }
    fr[0][c] = sum.x / m.vol[c];
    fr[1][c] = sum.y / m.vol[c];
    fr[2][c] = sum.z / m.vol[c];
  }
  for (int d = 0; d < 3; ++d) {
    ffr[d].resize(nf);
    Interpolate(m, fr[d].data(), ffr[d].data(), d0);
  }
  Vec ffmu(nf);
  Interpolate(m, o->mu_s.data(), ffmu.data(), d0);

  // pressure gradient and its face values
  Vec ffp(nf), gp[3], ffgp[3];
  Interpolate(m, o->p_ip.data(), ffp.data(), MakeBc(BC_EXTRAP));
  for (int d = 0; d < 3; ++d) gp[d].resize(nc);
  {
    double* gg[3] = {gp[0].data(), gp[1].data(), gp[2].data()};
    Gradient(m, ffp.data(), gg);
  }
  for (int d = 0; d < 3; ++d) {
    ffgp[d].resize(nf);
    Interpolate(m, gp[d].data(), ffgp[d].data(), d0);
  }

  // explicit viscosity (fluid.hpp:835-854)
  Vec fo[3];
  for (int e = 0; e < 3; ++e) fo[e].assign(nc, 0.);
  {
    Vec ff(nf), g[3], gf[3];
    for (int e = 0; e < 3; ++e) {
      g[e].resize(nc);
      gf[e].resize(nf);
    }
    for (int d = 0; d < 3; ++d) {
      Interpolate(m, o->u_ic[d].data(), ff.data(), VelBc(o, d));
      double* gg[3] = {g[0].data(), g[1].data(), g[2].data()};
      Gradient(m, ff.data(), gg);
      for (int e = 0; e < 3; ++e) Interpolate(m, g[e].data(), gf[e].data(), d0);
      for (size_t c = 0; c < nc; ++c) {
        V3 sum{0, 0, 0};
        for (int q = 0; q < 6; ++q) {
          const size_t fq = m.Face(c, q);
          const double w = ffmu[fq] * m.OutSurf(c, q)[d];
          sum.x += gf[0][fq] * w;
          sum.y += gf[1][fq] * w;
          sum.z += gf[2][fq] * w;
        }
        fo[0][c] += sum.x / m.vol[c];
        fo[1][c] += sum.y / m.vol[c];
        fo[2][c] += sum.z / m.vol[c];
      }
    }
  }
  // force assembly (fluid.hpp:857-870), zero sources
  for (size_t c = 0; c < nc; ++c) {
    const double s = o->rho_s[c] * 0. - 0.;
    for (int d = 0; d < 3; ++d) fo[d][c] += ((gp[d][c] * (-1.) + fr[d][c]) + 0.) + o->u_ic[d][c] * s;
  }

  // momentum rows (conv_diff.hpp:130-251): coefficients shared by the
  // components, constants per component
  Sys A;
  A.Resize(m.nx, m.ny, m.nz);
  Vec kk[3];
  int vel_iters[3] = {0, 0, 0};
  const Vec& probe = o->fl_ip;
  for (int d = 0; d < 3; ++d) {
    o->u_ip[d] = o->u_ic[d];
    const Vec& up = o->u_ip[d];
    Vec ff(nf), g[3];
    for (int e = 0; e < 3; ++e) g[e].resize(nc);
    const SideBc vbc = VelBc(o, d);
    Interpolate(m, up.data(), ff.data(), vbc);
    {
      double* gg[3] = {g[0].data(), g[1].data(), g[2].data()};
      Gradient(m, ff.data(), gg);
    }
    Vec cam(nf), cap(nf), cak(nf), dcm(nf), dcp(nf), dck(nf);
    for (size_t f = 0; f < nf; ++f) {
      const double F = probe[f], mu = ffmu[f], A_ = m.area[f];
      if (m.Inner(f)) {
        const long cm = m.fcm[f], cp = m.fcp[f];
        double am, ap, ak;
        if (F > 1e-10) {
          am = 1.;
          ap = 0.;
          ak = -Dot(V3{g[0][cm], g[1][cm], g[2][cm]}, m.to0[f]);
        } else if (F < -1e-10) {
          am = 0.;
          ap = 1.;
          ak = -Dot(V3{g[0][cp], g[1][cp], g[2][cp]}, m.to1[f]);
        } else {
          am = 0.5;
          ap = 0.5;
          ak = 0.;
        }
        cam[f] = am * F;
        cap[f] = ap * F;
        cak[f] = ak * F;
        const double al = 1. / Norm(Sub(m.to1[f], m.to0[f]));
        dcm[f] = ((-al) * (-mu)) * A_;
        dcp[f] = (al * (-mu)) * A_;
        dck[f] = (0. * (-mu)) * A_;
      } else {
        const int id = m.ValidId(f);
        const double val = vbc.val[m.side[f]];
        cam[f] = cap[f] = 0.;
        cak[f] = val * F;
        const double factor = (id == 0 ? 1. : -1.);
        const double al = 1. / Norm(m.ToCell(f, id)) * factor;
        const double dcc = ((-al) * (-mu)) * A_;
        dcm[f] = id == 0 ? dcc : 0.;
        dcp[f] = id == 0 ? 0. : dcc;
        dck[f] = ((al * val) * (-mu)) * A_;
      }
    }
    const std::vector<double>& co = o->bdf;
    const double relax = cfg.velocity_relaxation_factor;
    kk[d].assign(nc, 0.);
    for (size_t c = 0; c < nc; ++c) {
      double Cc = 0., Ck = 0., Dc = 0., Dk = 0.;
      double Cnb[6] = {0, 0, 0, 0, 0, 0}, Dnb[6] = {0, 0, 0, 0, 0, 0};
      for (int q = 0; q < 6; ++q) {
        const size_t fq = m.Face(c, q);
        const double o_ = Mesh::Out(q);
        const bool self_is_m = (q % 2 == 1);  // E, N, F faces: c is cm
        const double cself = self_is_m ? cam[fq] : cap[fq];
        const double cnb = self_is_m ? cap[fq] : cam[fq];
        const double dself = self_is_m ? dcm[fq] : dcp[fq];
        const double dnb = self_is_m ? dcp[fq] : dcm[fq];
        if (m.Inner(fq)) {
          Cc += cself * o_;
          Cnb[q] = cnb * o_;
          Dnb[q] = dnb * o_;
        }
        Dc += dself * o_;
        Ck += cak[fq] * o_;
        Dk += dck[fq] * o_;
      }
      const double V = m.vol[c], rho = o->rho_s[c];
      const double cen = (Dc / V) + ((Cc / V + co[2]) * rho);
      double nb[6];
      for (int q = 0; q < 6; ++q) nb[q] = (Dnb[q] / V) + ((Cnb[q] / V) * rho);
      const double unst = co[0] * o->u_tp[d][c] + co[1] * o->u_tc[d][c];
      const double kon = ((Dk / V) + ((Ck / V + unst) * rho)) + (-fo[d][c]);
      // row terms in index order; delta form: constant = row at u_prev
      double row[7];
      for (int q = 0; q < 6; ++q) row[kFaceTerm[q]] = nb[q];
      row[3] = cen;
      double ev = kon;
      for (int q = 0; q < 7; ++q)
        if (A.Has(c, q)) ev += up[long(c) + A.Off(q)] * row[q];
      kk[d][c] = ev;
      if (d == 0) {
        for (int q = 0; q < 7; ++q) A.t[q][c] = (q == 3) ? cen / relax : (A.Has(c, q) ? row[q] : 0.);
      }
    }
  }
  for (int d = 0; d < 3; ++d) {
    Vec corr;
    const SolveResult r = RunSolver(o, cfg.linear_solver_velocity, A, kk[d], corr);
    vel_iters[d] = r.iters;
    for (size_t c = 0; c < nc; ++c) o->u_ic[d][c] = o->u_ip[d][c] + corr[c];
  }
  o->cd_iter += 1;

  // diagonal coefficient (fluid.hpp:876-887): mean over the components
  Vec a_p(nc);
  for (size_t c = 0; c < nc; ++c) {
    const double cs = A.CoeffSum(c);
    double sum = 0.;
    sum += cs;
    sum += cs;
    sum += cs;
    a_p[c] = sum / 3.;
  }
  Vec ffd(nf);
  Interpolate(m, a_p.data(), ffd.data(), MakeBc(BC_NONE));
  // Rhie-Chow face fluxes (fluid.hpp:888-970)
  Vec fu[3], fstar(nf), fcoef(nf, 0.);
  for (int d = 0; d < 3; ++d) {
    fu[d].resize(nf);
    Interpolate(m, o->u_ic[d].data(), fu[d].data(), VelBc(o, d));
  }
  const double rc = cfg.rhie_chow_factor;
  const V3 mv{0., 0., 0.};
  for (size_t f = 0; f < nf; ++f) {
    const double vfi = Dot(V3{fu[0][f], fu[1][f], fu[2][f]}, m.surf[f]);
    if (m.Inner(f)) {
      const long cm = m.fcm[f], cp = m.fcp[f];
      const V3 dd = Sub(m.to1[f], m.to0[f]);
      const double wide = Dot(V3{ffgp[0][f] - ffr[0][f], ffgp[1][f] - ffr[1][f],
                                 ffgp[2][f] - ffr[2][f]}, m.surf[f]);
      const double compact = (o->p_ip[cp] - o->p_ip[cm]) / Norm(dd) * m.area[f] -
                             Dot(V3{ffe[0][f], ffe[1][f], ffe[2][f]}, m.surf[f]);
      fstar[f] = vfi + rc * (wide - compact) / ffd[f] + 0;
    } else {
      fstar[f] = vfi;
    }
  }
  for (size_t f = 0; f < nf; ++f) fstar[f] -= Dot(mv, m.surf[f]);
  for (size_t f = 0; f < nf; ++f)
    if (m.Inner(f)) fcoef[f] = -m.area[f] / (Norm(Sub(m.to1[f], m.to0[f])) * ffd[f]);
  // pressure-correction rows (fluid.hpp:972-1027)
  Sys P;
  P.Resize(m.nx, m.ny, m.nz);
  for (size_t c = 0; c < nc; ++c) {
    double cen = 0., kon = 0.;
    double nb[6] = {0, 0, 0, 0, 0, 0};
    for (int q = 0; q < 6; ++q) {
      const size_t fq = m.Face(c, q);
      const double o_ = Mesh::Out(q);
      const bool self_is_m = (q % 2 == 1);
      if (m.Inner(fq)) {
        const double cself = self_is_m ? -fcoef[fq] : fcoef[fq];
        const double cnb = self_is_m ? fcoef[fq] : -fcoef[fq];
        cen += cself * o_;
        nb[q] = cnb * o_;
      }
      kon += fstar[fq] * o_;
    }
    kon += -(0. * m.vol[c]);
    for (int q = 0; q < 6; ++q) P.t[kFaceTerm[q]][c] = nb[q];
    P.t[3][c] = cen;
    P.k[c] = kon;
  }
  Vec pc;
  const SolveResult pr = RunSolver(o, cfg.linear_solver_pressure, P, P.k, pc);
  // corrections (fluid.hpp:1033-1057)
  for (size_t c = 0; c < nc; ++c) o->p_ic[c] = o->p_ip[c] + cfg.pressure_relaxation_factor * pc[c];
  {
    Vec ffpc(nf), g[3];
    for (int e = 0; e < 3; ++e) g[e].resize(nc);
    Interpolate(m, pc.data(), ffpc.data(), MakeBc(BC_EXTRAP));
    double* gg[3] = {g[0].data(), g[1].data(), g[2].data()};
    Gradient(m, ffpc.data(), gg);
    for (size_t c = 0; c < nc; ++c)
      for (int d = 0; d < 3; ++d) o->u_ic[d][c] += g[d][c] / (-a_p[c]);
  }
  for (size_t f = 0; f < nf; ++f) {
    double v = fstar[f];
    if (m.Inner(f)) {
      v += pc[m.fcm[f]] * (-fcoef[f]);
      v += pc[m.fcp[f]] * fcoef[f];
    }
    o->fl_ic[f] = v;
  }
  o->fluid_iter += 1;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->iteration = o->fluid_iter;
    stats->convergence_indicator = ConvergenceIndicator(o);
    stats->velocity_solve_iters[0] = vel_iters[0];
    stats->velocity_solve_iters[1] = vel_iters[1];
    stats->pressure_solve_iters = pr.iters;
    stats->pressure_solve_diff = pr.diff;
    stats->pressure_converged = pr.converged ? 1 : 0;
  }
  return HG_OK;
}

bool IsConverged(const or3_ctx* o) {
  const double ind = o->cd_iter == 0 ? 1. : ConvergenceIndicator(o);
  return o->fluid_iter >= o->cfg.num_iterations_limit || ind < o->cfg.convergence_tolerance;
}

int FinishStep(or3_ctx* o) {
  o->p_tp = o->p_tc;
  o->fl_tp = o->fl_tc;
  o->p_tc = o->p_ic;
  o->fl_tc = o->fl_ic;
  for (double v : o->p_tc)
    if (!(v * 0. == 0.)) return HG_ERR_NAN;
  for (int d = 0; d < 3; ++d) {
    o->u_tp[d] = o->u_tc[d];
    o->u_tc[d] = o->u_ic[d];
    for (double v : o->u_tc[d])
      if (!(v * 0. == 0.)) return HG_ERR_NAN;
  }
  o->t_fluid += o->dt;
  return HG_OK;
}

// AdvectionSolverMultiExplicit (advection.hpp:417-545): unsplit Superbee,
// zero-derivative boundaries, no sharpening, no sources
void AdvectSubstep(or3_ctx* o) {
  const Mesh& m = o->m;
  const size_t nc = m.nc, nf = m.nf;
  for (int i = 0; i < o->np; ++i) {
    o->pd_tp[i] = o->pd_tc[i];
    o->pd_ic[i] = o->pd_tp[i];
  }
  const Vec& flux = o->fl_ic;
  const SideBc d0 = MakeBc(BC_DERIV, 0.);
  for (int i = 0; i < o->np; ++i) {
    Vec& curr = o->pd_ic[i];
    Vec vf(nf);
    for (size_t f = 0; f < nf; ++f) vf[f] = flux[f] + 0.;
    Vec uf(nf), g[3], fu(nf);
    for (int e = 0; e < 3; ++e) g[e].resize(nc);
    Interpolate(m, curr.data(), uf.data(), d0);
    double* gg[3] = {g[0].data(), g[1].data(), g[2].data()};
    Gradient(m, uf.data(), gg);
    for (size_t f = 0; f < nf; ++f) {
      if (m.Inner(f)) {
        const long P = m.fcm[f], E = m.fcp[f];
        const Vec& u = curr;
        if (vf[f] > 1e-8) {
          fu[f] = u[P] + 0.5 * Superbee(u[E] - u[P], -4. * Dot(V3{g[0][P], g[1][P], g[2][P]},
                                                               m.to0[f]) - (u[E] - u[P]));
        } else if (vf[f] < -1e-8) {
          fu[f] = u[E] - 0.5 * Superbee(u[E] - u[P], 4. * Dot(V3{g[0][E], g[1][E], g[2][E]},
                                                              m.to1[f]) - (u[E] - u[P]));
        } else {
          fu[f] = 0.5 * (u[P] + u[E]);
        }
      } else {
        const int id = m.ValidId(f);
        const long cc = id == 0 ? m.fcm[f] : m.fcp[f];
        const double factor = (id == 0 ? 1. : -1.);
        const double al = Norm(m.ToCell(f, id)) * factor;
        fu[f] = curr[cc] + 0. * al;
      }
    }
    const double dt = o->dt_adv;
    for (size_t c = 0; c < nc; ++c) {
      double fs = 0.;
      for (int q = 0; q < 6; ++q) {
        const size_t fq = m.Face(c, q);
        fs += fu[fq] * vf[fq] * Mesh::Out(q);
      }
      curr[c] += -dt / m.vol[c] * fs;
    }
    for (size_t c = 0; c < nc; ++c) {
      curr[c] += dt * 0. * 0.;
      double sh = 0.;
      for (int q = 0; q < 6; ++q) sh += Mesh::Out(q) * 0.;
      curr[c] += dt * sh / m.vol[c];
    }
  }
  for (int i = 0; i < o->np; ++i) o->pd_tc[i] = o->pd_ic[i];
  o->t_adv += o->dt_adv;
}

Vec* FieldPtr(or3_ctx* o, int field, bool write) {
  if (field == HG_FIELD_VELOCITY_X) return &o->u_tc[0];
  if (field == HG_FIELD_VELOCITY_Y) return &o->u_tc[1];
  if (field == HG_FIELD_VELOCITY_Z) return &o->u_tc[2];
  if (field == HG_FIELD_PRESSURE) return &o->p_tc;
  if (field == HG_FIELD_VOLUME_FLUX) return &o->fl_tc;
  if (field >= HG_FIELD_PARTIAL_DENSITY_0 && field < HG_FIELD_PARTIAL_DENSITY_0 + o->np)
    return &o->pd_tc[field - HG_FIELD_PARTIAL_DENSITY_0];
  if (write) return nullptr;
  if (field >= HG_FIELD_VOLUME_FRACTION_0 && field < HG_FIELD_VOLUME_FRACTION_0 + o->np)
    return &o->alpha[field - HG_FIELD_VOLUME_FRACTION_0];
  if (field == HG_FIELD_DENSITY) return &o->rho;
  if (field == HG_FIELD_DENSITY_SMOOTH) return &o->rho_s;
  if (field == HG_FIELD_VISCOSITY_SMOOTH) return &o->mu_s;
  if (field == HG_FIELD_FORCE_X) return &o->f[0];
  if (field == HG_FIELD_FORCE_Y) return &o->f[1];
  if (field == HG_FIELD_FORCE_Z) return &o->f[2];
  return nullptr;
}

}  // namespace

extern "C" {

int or3_create(const hg3_config* cfg3, or3_ctx** out, char* err, size_t errlen) {
  auto fail = [&](const char* msg) {
    if (err && errlen) std::snprintf(err, errlen, "%s", msg);
    return int(HG_ERR_UNSUPPORTED);
  };
  const hg_config& cfg = cfg3->base;
  if (cfg.nx < 2 || cfg.ny < 2 || cfg3->nz < 2) return fail("Nx, Ny, Nz must be >= 2");
  if (cfg.num_phases < 1 || cfg.num_phases > 2) return fail("num_phases 1 or 2");
  for (int s = 0; s < 6; ++s)
    if (cfg3->bc_kind[s] != HG_BC_WALL) return fail("walls only");
  or3_ctx* o = new or3_ctx();
  o->cfg3 = *cfg3;
  o->cfg = cfg;
  o->np = cfg.num_phases;
  const double a[3] = {cfg.a[0], cfg.a[1], cfg3->a_z}, b[3] = {cfg.b[0], cfg.b[1], cfg3->b_z};
  o->m.Init(cfg.nx, cfg.ny, cfg3->nz, a, b);
  const Mesh& m = o->m;
  const size_t nc = m.nc;
  o->g[0] = cfg.gravity[0];
  o->g[1] = cfg.gravity[1];
  o->g[2] = cfg3->gravity_z;
  o->force[0] = cfg.force[0];
  o->force[1] = cfg.force[1];
  o->force[2] = cfg3->force_z;
  o->dt = cfg.dt;
  o->dt_adv = cfg.dt * cfg.advection_dt_factor;
  o->bdf = DerivCoeffs(0., {-2. * o->dt, -o->dt, 0.}, cfg.time_second_order ? 0 : 1);
  const double iv[3] = {cfg.initial_velocity[0], cfg.initial_velocity[1],
                        cfg3->initial_velocity_z};
  for (int d = 0; d < 3; ++d) {
    o->u_tc[d].assign(nc, iv[d]);
    o->u_tp[d] = o->u_tc[d];
  }
  o->p_tc.assign(nc, 0.);
  o->p_tp = o->p_tc;
  InitialFlux(o);
  // initial partial densities (hydro2d.hpp:494-566): phase 1 in A1-B1
  for (int i = 0; i < o->np; ++i) o->pd_tc[i].assign(nc, cfg.density[i] * cfg.initial_volume_fraction[i]);
  const double lo[3] = {cfg.a1[0], cfg.a1[1], cfg3->a1_z}, hi[3] = {cfg.b1[0], cfg.b1[1], cfg3->b1_z};
  if (o->np >= 2) {
    for (size_t c = 0; c < nc; ++c) {
      const V3 x = m.cc[c];
      bool in = true;
      for (int d = 0; d < 3; ++d) in = in && !(x[d] < lo[d] || hi[d] < x[d]);
      if (in) o->pd_tc[1][c] = cfg.density[1];
    }
    Smooth(m, o->pd_tc[1], cfg.initial_volume_fraction_smooth_times);
  }
  for (size_t c = 0; c < nc; ++c) {
    double vs = 0.;
    for (int i = 1; i < o->np; ++i) vs += o->pd_tc[i][c] / cfg.density[i];
    o->pd_tc[0][c] = (1. - vs) * cfg.density[0];
  }
  for (int i = 0; i < o->np; ++i) {
    o->pd_tp[i] = o->pd_tc[i];
    o->alpha[i].assign(nc, 0.);
  }
  UpdateProperties(o);
  *out = o;
  return HG_OK;
}

void or3_destroy(or3_ctx* o) { delete o; }
int or3_num_cells(const or3_ctx* o) { return int(o->m.nc); }
int or3_num_faces(const or3_ctx* o) { return int(o->m.nf); }

int or3_get_field(or3_ctx* o, int field, double* host, size_t n) {
  Vec* v = FieldPtr(o, field, false);
  if (!v || v->size() != n) return HG_ERR_BAD_ARG;
  std::copy(v->begin(), v->end(), host);
  return HG_OK;
}

int or3_set_field(or3_ctx* o, int field, const double* host, size_t n) {
  Vec* v = FieldPtr(o, field, true);
  if (!v || v->size() != n) return HG_ERR_BAD_ARG;
  std::copy(host, host + n, v->begin());
  if (field == HG_FIELD_VELOCITY_X) o->u_tp[0] = *v;
  if (field == HG_FIELD_VELOCITY_Y) o->u_tp[1] = *v;
  if (field == HG_FIELD_VELOCITY_Z) o->u_tp[2] = *v;
  if (field == HG_FIELD_PRESSURE) o->p_tp = *v;
  if (field == HG_FIELD_VOLUME_FLUX) o->fl_tp = *v;
  if (field >= HG_FIELD_PARTIAL_DENSITY_0 && field < HG_FIELD_PARTIAL_DENSITY_0 + o->np)
    o->pd_tp[field - HG_FIELD_PARTIAL_DENSITY_0] = *v;
  return HG_OK;
}

// hydro::step (hydro2d.hpp:1531-1621)
int or3_step(or3_ctx* o, hg_step_stats* stats) {
  hg_step_stats s{};
  int rc = StartStep(o);
  if (rc) return rc;
  s.pressure_converged = 1;
  if (o->cfg.fluid_enable) {
    while (!IsConverged(o)) {
      hg_iter_stats is;
      rc = Iterate(o, &is);
      if (rc) return rc;
      s.pressure_solve_iters_total += is.pressure_solve_iters;
      s.convergence_indicator = is.convergence_indicator;
      s.pressure_solve_diff_max = std::max(s.pressure_solve_diff_max, is.pressure_solve_diff);
      s.pressure_converged = s.pressure_converged && is.pressure_converged;
    }
  }
  s.simple_iterations = o->fluid_iter;
  rc = FinishStep(o);
  if (rc) return rc;
  if (o->cfg.advection_enable) {
    while (o->t_adv < o->t_fluid - 0.5 * o->dt_adv) {
      AdvectSubstep(o);
      s.advection_substeps += 1;
    }
  }
  UpdateProperties(o);
  s.time = o->t_fluid;
  s.dt = o->dt;
  if (stats) *stats = s;
  return HG_OK;
}

}  // extern "C"
