// hydro_oracle.cpp -- CPU oracle of the hydro2d time step (TEST INFRASTRUCTURE).
//
// A restatement of the reference algorithm on structure-of-arrays fields.
// The floating-point operations of every value follow the reference's own
// order (Expression merge order, Vect component order, left-to-right
// products), so with -ffp-contract=off the oracle reproduces the reference
// bit for bit; each function names the reference lines it restates.
// Loops that write one element each run under OpenMP (as the reference's 14
// `omp parallel for` sites do); solvers and reductions stay serial.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
// this library.  The product (hydro_amd/libhydrogpu.so) does not.

#include "hydro_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace {

using Vec = std::vector<double>;

struct V2 {
  double x, y;
};

// Vect::dot, sqrnorm, norm (common/vect.hpp:173-192): sums start from 0.
inline double Dot(V2 a, V2 b) {
  double s = 0;
  s += a.x * b.x;
  s += a.y * b.y;
  return s;
}
inline double Norm(V2 a) {
  double s = 0;
  s += a.x * a.x;
  s += a.y * a.y;
  return std::sqrt(s);
}

enum BcKind { BC_NONE = 0, BC_VALUE, BC_DERIV, BC_EXTRAP };

// One boundary condition per domain side (left, right, bottom, top).
struct SideBc {
  BcKind kind[4];
  double val[4];
  // per-face values of BC_VALUE sides (indexed by face), when they vary
  // along a side (outlets, inlet partial densities)
  const double* fval = nullptr;
};

SideBc MakeBc(BcKind k, double v = 0.) {
  SideBc b;
  for (int s = 0; s < 4; ++s) {
    b.kind[s] = k;
    b.val[s] = v;
  }
  return b;
}

// Uniform structured 2D mesh with the reference's geometry formulas
// (mesh.hpp:722-736 InitUniformMesh; mesh2d.hpp:215-245 CalcCenter/Area/
// Surface/Volume; mesh2d.hpp:392-404 vect to cell).
struct Mesh {
  int nx = 0, ny = 0;
  size_t nc = 0, nfx = 0, nf = 0;
  Vec xn, yn;
  std::vector<V2> cc;
  Vec vol;
  std::vector<V2> fc, surf, normal, to0, to1;
  Vec area;
  std::vector<long> fcm, fcp;  // face neighbour cells (-1: none)

  size_t FW(int i, int j) const { return size_t(i) + size_t(nx + 1) * j; }
  size_t FE(int i, int j) const { return size_t(i + 1) + size_t(nx + 1) * j; }
  size_t FS(int i, int j) const { return nfx + size_t(i) + size_t(nx) * j; }
  size_t FN(int i, int j) const { return nfx + size_t(i) + size_t(nx) * (j + 1); }
  size_t Face(size_t c, int k) const {
    int i = int(c % nx), j = int(c / nx);
    switch (k) {
      case 0: return FW(i, j);
      case 1: return FE(i, j);
      case 2: return FS(i, j);
      default: return FN(i, j);
    }
  }
  // GetOutwardFactor (mesh2d.hpp:118-129): W,S -> -1; E,N -> +1.
  static double Out(int k) { return (k == 0 || k == 2) ? -1. : 1.; }
  V2 OutSurf(size_t c, int k) const {
    V2 s = surf[Face(c, k)];
    double f = Out(k);
    return V2{s.x * f, s.y * f};
  }
  bool Inner(size_t f) const { return fcm[f] >= 0 && fcp[f] >= 0; }
  int ValidId(size_t f) const { return fcm[f] >= 0 ? 0 : 1; }
  int Side(size_t f) const {
    if (f < nfx) return (f % (nx + 1)) == 0 ? 0 : 1;
    return ((f - nfx) / nx) == 0 ? 2 : 3;
  }
  V2 ToCell(size_t f, int id) const { return id == 0 ? to0[f] : to1[f]; }

  void Init(int nx_, int ny_, const double a[2], const double b[2]) {
    nx = nx_;
    ny = ny_;
    nc = size_t(nx) * ny;
    nfx = size_t(nx + 1) * ny;
    nf = nfx + size_t(nx) * (ny + 1);
    // Nodes: lb + (midx / size) * (rt - lb), component-wise.
    xn.resize(nx + 1);
    yn.resize(ny + 1);
    double dx = b[0] - a[0], dy = b[1] - a[1];
    for (int i = 0; i <= nx; ++i) xn[i] = a[0] + (double(i) / double(nx)) * dx;
    for (int j = 0; j <= ny; ++j) yn[j] = a[1] + (double(j) / double(ny)) * dy;
    fc.resize(nf);
    surf.resize(nf);
    normal.resize(nf);
    area.resize(nf);
    to0.assign(nf, V2{0, 0});
    to1.assign(nf, V2{0, 0});
    fcm.resize(nf);
    fcp.resize(nf);
    auto face_geom = [&](size_t f, V2 n0, V2 n1) {
      V2 c{0, 0};
      c.x += n0.x;
      c.y += n0.y;
      c.x += n1.x;
      c.y += n1.y;
      fc[f] = V2{c.x / 2., c.y / 2.};
      area[f] = Norm(V2{n0.x - n1.x, n0.y - n1.y});  // n1.dist(n0)
      V2 v{n1.x - n0.x, n1.y - n0.y};
      surf[f] = V2{v.y, -v.x};
    };
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        size_t f = FW(i, j);
        face_geom(f, V2{xn[i], yn[j]}, V2{xn[i], yn[j + 1]});
        fcm[f] = i > 0 ? long(i - 1) + long(nx) * j : -1;
        fcp[f] = i < nx ? long(i) + long(nx) * j : -1;
      }
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t f = nfx + size_t(i) + size_t(nx) * j;
        face_geom(f, V2{xn[i + 1], yn[j]}, V2{xn[i], yn[j]});
        fcm[f] = j > 0 ? long(i) + long(nx) * (j - 1) : -1;
        fcp[f] = j < ny ? long(i) + long(nx) * j : -1;
      }
    cc.resize(nc);
    vol.resize(nc);
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t c = size_t(i) + size_t(nx) * j;
        V2 s{0, 0};
        V2 nd[4] = {{xn[i], yn[j]}, {xn[i + 1], yn[j]}, {xn[i + 1], yn[j + 1]}, {xn[i], yn[j + 1]}};
        for (int k = 0; k < 4; ++k) {
          s.x += nd[k].x;
          s.y += nd[k].y;
        }
        cc[c] = V2{s.x / 4., s.y / 4.};
      }
    for (size_t c = 0; c < nc; ++c) {
      double r = 0.;
      for (int k = 0; k < 4; ++k) r += fc[Face(c, k)].x * OutSurf(c, k).x;
      vol[c] = r;
    }
    for (size_t f = 0; f < nf; ++f) {
      if (fcm[f] >= 0) to0[f] = V2{cc[fcm[f]].x - fc[f].x, cc[fcm[f]].y - fc[f].y};
      if (fcp[f] >= 0) to1[f] = V2{cc[fcp[f]].x - fc[f].x, cc[fcp[f]].y - fc[f].y};
      normal[f] = V2{surf[f].x / area[f], surf[f].y / area[f]};
    }
  }
};

// Interpolate cell -> face (solver.hpp:392-470) for one scalar field (a Vect
// field is interpolated component by component, which is what the reference's
// per-component Vect arithmetic amounts to).
void Interpolate(const Mesh& m, const double* u, double* res, const SideBc& bc) {
  const long nf = long(m.nf);
#pragma omp parallel for
  for (long f = 0; f < nf; ++f) {
    if (m.Inner(f))
      res[f] = u[m.fcm[f]] * (1. - 0.5) + u[m.fcp[f]] * 0.5;
    else
      res[f] = 0.;
  }
  for (size_t f = 0; f < m.nf; ++f) {
    if (m.Inner(f)) continue;
    int side = m.Side(f);
    BcKind kind = bc.kind[side];
    if (kind == BC_NONE) continue;
    int id = m.ValidId(f);
    size_t c = size_t(id == 0 ? m.fcm[f] : m.fcp[f]);
    double factor = (id == 0 ? 1. : -1.);
    if (kind == BC_VALUE) {
      res[f] = bc.fval ? bc.fval[f] : bc.val[side];
    } else if (kind == BC_DERIV) {
      double alpha = Norm(m.ToCell(f, id)) * factor;
      res[f] = u[c] + bc.val[side] * alpha;
    } else {
      V2 nrm{m.normal[f].x * factor, m.normal[f].y * factor};
      double dist = Norm(m.ToCell(f, id));
      double nom = u[c] / dist;
      double den = 1. / dist;
      double volume = m.vol[c];
      for (int k = 0; k < 4; ++k) {
        size_t nface = m.Face(c, k);
        if (nface == f)
          den -= Dot(m.OutSurf(c, k), nrm) / volume;
        else
          nom += res[nface] * Dot(m.OutSurf(c, k), nrm) / volume;
      }
      res[f] = nom / den;
    }
  }
}

// Gradient (solver.hpp:658-677).
void Gradient(const Mesh& m, const double* uf, double* gx, double* gy) {
  const long nc = long(m.nc);
#pragma omp parallel for
  for (long c = 0; c < nc; ++c) {
    V2 s{0, 0};
    for (int k = 0; k < 4; ++k) {
      V2 o = m.OutSurf(c, k);
      double u = uf[m.Face(c, k)];
      s.x += o.x * u;
      s.y += o.y * u;
    }
    gx[c] = s.x / m.vol[c];
    gy[c] = s.y / m.vol[c];
  }
}

// GetSmoothField (solver.hpp:636-656) with Average (solver.hpp:621-634).
void Smooth(const Mesh& m, Vec& u, int repeat) {
  SideBc bc = MakeBc(BC_DERIV, 0.);
  Vec uf(m.nf);
  for (int r = 0; r < repeat; ++r) {
    Interpolate(m, u.data(), uf.data(), bc);
    const long nc = long(m.nc);
#pragma omp parallel for
    for (long c = 0; c < nc; ++c) {
      double s = 0;
      for (int k = 0; k < 4; ++k) s += uf[m.Face(c, k)];
      u[c] = s / 4.;
    }
  }
}

// Superbee limiter (solver.hpp:550-558).
double Superbee(double p, double q) {
  if (p > 0. && q > 0.) return std::max(std::min(2 * p, q), std::min(p, 2 * q));
  if (p < 0. && q < 0.) return -std::max(std::min(-2 * p, -q), std::min(-p, -2 * q));
  return 0.;
}

// GetDerivativeApproxCoeffs (solver.hpp:815-861).
std::vector<double> DerivCoeffs(double target, const std::vector<double>& args) {
  size_t size = args.size();
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

// ---- 5-point systems: row r = s x[r-nx] + w x[r-1] + c x[r] + e x[r+1] + n x[r+nx] + k
struct Sys {
  int nx = 0, ny = 0;
  Vec s, w, c, e, n, k;
  void Resize(int nx_, int ny_) {
    nx = nx_;
    ny = ny_;
    size_t N = size_t(nx) * ny;
    s.assign(N, 0.);
    w.assign(N, 0.);
    c.assign(N, 0.);
    e.assign(N, 0.);
    n.assign(N, 0.);
    k.assign(N, 0.);
  }
  size_t size() const { return size_t(nx) * ny; }
  bool HasS(size_t r) const { return r >= size_t(nx); }
  bool HasW(size_t r) const { return r % nx != 0; }
  bool HasE(size_t r) const { return r % nx != size_t(nx - 1); }
  bool HasN(size_t r) const { return r + nx < size(); }
  // Expression::Evaluate (linear.hpp:128-135): constant, then terms in
  // index order S, W, C, E, N.
  double Eval(size_t r, const double* x) const {
    double v = k[r];
    if (HasS(r)) v += x[r - nx] * s[r];
    if (HasW(r)) v += x[r - 1] * w[r];
    v += x[r] * c[r];
    if (HasE(r)) v += x[r + 1] * e[r];
    if (HasN(r)) v += x[r + nx] * n[r];
    return v;
  }
  // Expression::CoeffSum (linear.hpp:137-142).
  double CoeffSum(size_t r) const {
    double v = 0.;
    if (HasS(r)) v += s[r];
    if (HasW(r)) v += w[r];
    v += c[r];
    if (HasE(r)) v += e[r];
    if (HasN(r)) v += n[r];
    return v;
  }
};

struct SolveResult {
  int iters = 0;
  double diff = 0.;
  bool converged = true;
};

// LuDecomposition::Solve (linear.hpp:527-567): one symmetric GS sweep from 0.
void SolveLu(const Sys& a, const Vec& k, Vec& x) {
  size_t N = a.size();
  x.assign(N, 0.);
  for (size_t i = 0; i < N; ++i) {
    double sum = 0;
    if (a.HasS(i)) sum += a.s[i] * x[i - a.nx];
    if (a.HasW(i)) sum += a.w[i] * x[i - 1];
    x[i] = (-k[i] - sum) / a.c[i];
  }
  for (size_t i = N; i > 0;) {
    --i;
    double sum = 0;
    if (a.HasN(i)) sum += a.n[i] * x[i + a.nx];
    if (a.HasE(i)) sum += a.e[i] * x[i + 1];
    x[i] -= sum / a.c[i];
  }
}

// "lu_rb" (new, no reference counterpart): the symmetric GS sweep of SolveLu
// taken in red-black order -- red cells ((i+j) even) first, then black.
// Forward: red x = -k/c; black x = (-k - sum_nb)/c with neighbours summed in
// S, W, E, N order.  Backward: black rows have no later neighbour; red
// x -= (sum over N, E, W, S) / c.
void SolveLuRb(const Sys& a, const Vec& k, Vec& x) {
  size_t N = a.size();
  x.assign(N, 0.);
  const int nx = a.nx;
  auto red = [nx](size_t r) { return ((r % nx) + (r / nx)) % 2 == 0; };
  for (size_t i = 0; i < N; ++i)
    if (red(i)) x[i] = (-k[i] - 0.) / a.c[i];
  for (size_t i = 0; i < N; ++i) {
    if (red(i)) continue;
    double sum = 0;
    if (a.HasS(i)) sum += a.s[i] * x[i - nx];
    if (a.HasW(i)) sum += a.w[i] * x[i - 1];
    if (a.HasE(i)) sum += a.e[i] * x[i + 1];
    if (a.HasN(i)) sum += a.n[i] * x[i + nx];
    x[i] = (-k[i] - sum) / a.c[i];
  }
  for (size_t i = 0; i < N; ++i) {
    if (!red(i)) continue;
    double sum = 0;
    if (a.HasN(i)) sum += a.n[i] * x[i + nx];
    if (a.HasE(i)) sum += a.e[i] * x[i + 1];
    if (a.HasW(i)) sum += a.w[i] * x[i - 1];
    if (a.HasS(i)) sum += a.s[i] * x[i - nx];
    x[i] -= sum / a.c[i];
  }
}

// GaussSeidel::Solve (linear.hpp:685-715): lexicographic SOR from 0.
SolveResult SolveGs(const Sys& a, const Vec& k, Vec& x, double tol, int limit, double omega) {
  size_t N = a.size();
  x.assign(N, 0.);
  size_t iter = 0;
  double diff = 0.;
  do {
    diff = 0.;
    for (size_t i = 0; i < N; ++i) {
      double sum = 0.;
      if (a.HasS(i)) sum += a.s[i] * x[i - a.nx];
      if (a.HasW(i)) sum += a.w[i] * x[i - 1];
      if (a.HasE(i)) sum += a.e[i] * x[i + 1];
      if (a.HasN(i)) sum += a.n[i] * x[i + a.nx];
      double value = -(k[i] + sum) / a.c[i];
      double corr = value - x[i];
      diff = std::max(diff, std::abs(corr));
      x[i] += corr * omega;
    }
  } while (diff > tol && iter++ < size_t(limit));
  SolveResult r;
  r.iters = int(iter);
  r.diff = diff;
  r.converged = !(diff > tol);
  return r;
}

// Jacobi::Solve (linear.hpp:750-783).
SolveResult SolveJacobi(const Sys& a, const Vec& k, Vec& x, double tol, int limit, double omega) {
  size_t N = a.size();
  x.assign(N, 0.);
  Vec next = x;
  size_t iter = 0;
  double diff = 0.;
  do {
    diff = 0.;
    for (size_t i = 0; i < N; ++i) {
      double sum = 0.;
      if (a.HasS(i)) sum += a.s[i] * x[i - a.nx];
      if (a.HasW(i)) sum += a.w[i] * x[i - 1];
      if (a.HasE(i)) sum += a.e[i] * x[i + 1];
      if (a.HasN(i)) sum += a.n[i] * x[i + a.nx];
      double value = -(k[i] + sum) / a.c[i];
      double corr = value - x[i];
      diff = std::max(diff, std::abs(corr));
      next[i] = x[i] + corr * omega;
    }
    std::swap(x, next);
  } while (diff > tol && iter++ < size_t(limit));
  SolveResult r;
  r.iters = int(iter);
  r.diff = diff;
  r.converged = !(diff > tol);
  return r;
}

// LuDecompositionRelaxed::Solve (linear.hpp:592-648).  The backward step
// reads `res`, not `corr`, exactly as the reference does.
SolveResult SolveLuRelaxed(const Sys& a, const Vec& k, Vec& res, double tol, int limit,
                           double relax) {
  size_t N = a.size();
  res.assign(N, 0.);
  Vec corr(N, 0.), f(k);
  size_t iter = 0;
  double diff = 0.;
  do {
    diff = 0.;
    for (size_t i = 0; i < N; ++i) {
      double sum = 0;
      if (a.HasS(i)) sum += a.s[i] * corr[i - a.nx];
      if (a.HasW(i)) sum += a.w[i] * corr[i - 1];
      corr[i] = (-f[i] - sum) / (a.c[i] + relax);
    }
    for (size_t i = N; i > 0;) {
      --i;
      double sum = 0;
      if (a.HasN(i)) sum += a.n[i] * res[i + a.nx];
      if (a.HasE(i)) sum += a.e[i] * res[i + 1];
      corr[i] -= sum / (a.c[i] + relax);
    }
    for (size_t i = 0; i < N; ++i) {
      res[i] += corr[i];
      diff = std::max(diff, std::abs(corr[i]));
    }
    for (size_t i = 0; i < N; ++i) {  // f = system evaluated at res
      double v = k[i];
      if (a.HasS(i)) v += res[i - a.nx] * a.s[i];
      if (a.HasW(i)) v += res[i - 1] * a.w[i];
      v += res[i] * a.c[i];
      if (a.HasE(i)) v += res[i + 1] * a.e[i];
      if (a.HasN(i)) v += res[i + a.nx] * a.n[i];
      f[i] = v;
    }
  } while (diff > tol && iter++ < size_t(limit));
  SolveResult r;
  r.iters = int(iter);
  r.diff = diff;
  r.converged = !(diff > tol);
  return r;
}

// A x for the 5-point rows without the constant: (A x)_r = sum coeff * x.
void Apply(const Sys& a, const double* x, double* y) {
  const long N = long(a.size());
#pragma omp parallel for
  for (long r = 0; r < N; ++r) {
    double v = a.c[r] * x[r];
    if (a.HasS(r)) v += a.s[r] * x[r - a.nx];
    if (a.HasW(r)) v += a.w[r] * x[r - 1];
    if (a.HasE(r)) v += a.e[r] * x[r + 1];
    if (a.HasN(r)) v += a.n[r] * x[r + a.nx];
    y[r] = v;
  }
}

double DotV(const Vec& a, const Vec& b) {
  double s = 0.;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

// "cg" (new): Jacobi-preconditioned CG from x = 0 on A x = b, b = -k,
// stopping at ||b - A x||_2 <= tol ||b||_2.  If `project` is set the system
// is the singular Neumann one: b is projected to zero mean first.
SolveResult SolveCg(const Sys& a, const Vec& k, Vec& x, double tol, int max_iters, bool project) {
  size_t N = a.size();
  Vec b(N);
  for (size_t i = 0; i < N; ++i) b[i] = -k[i];
  if (project) {
    double mean = 0.;
    for (double v : b) mean += v;
    mean /= double(N);
    for (double& v : b) v -= mean;
  }
  x.assign(N, 0.);
  Vec r = b, z(N), p(N), q(N);
  double bnorm = std::sqrt(DotV(b, b));
  SolveResult res;
  if (bnorm == 0.) {
    res.iters = 0;
    res.diff = 0.;
    return res;
  }
  for (size_t i = 0; i < N; ++i) z[i] = r[i] / a.c[i];
  p = z;
  double rz = DotV(r, z);
  int it = 0;
  double rel = 1.;
  while (it < max_iters) {
    Apply(a, p.data(), q.data());
    double alpha = rz / DotV(p, q);
    for (size_t i = 0; i < N; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    ++it;
    rel = std::sqrt(DotV(r, r)) / bnorm;
    if (rel <= tol) break;
    for (size_t i = 0; i < N; ++i) z[i] = r[i] / a.c[i];
    double rz2 = DotV(r, z);
    double beta = rz2 / rz;
    rz = rz2;
    for (size_t i = 0; i < N; ++i) p[i] = z[i] + beta * p[i];
  }
  // true residual
  Apply(a, x.data(), q.data());
  double t = 0.;
  for (size_t i = 0; i < N; ++i) t += (b[i] - q[i]) * (b[i] - q[i]);
  res.iters = it;
  res.diff = std::sqrt(t) / bnorm;
  res.converged = res.diff <= tol;
  return res;
}

struct Timers {
  std::map<std::string, double> total;
};

struct ScopedTimer {
  Timers* t;
  std::string name;
  std::chrono::steady_clock::time_point t0;
  ScopedTimer(Timers* tt, const char* n) : t(tt), name(n), t0(std::chrono::steady_clock::now()) {}
  ~ScopedTimer() {
    t->total[name] +=
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

}  // namespace

struct or_ctx {
  hg_config cfg;
  Mesh m;
  int np = 1;
  double dt = 0., dt_adv = 0.;
  double t_fluid = 0., t_adv = 0.;
  std::vector<double> bdf;  // unsteady coefficients
  // velocity components: time_curr, time_prev, iter_curr, iter_prev
  Vec u_tc[2], u_tp[2], u_ic[2], u_ip[2];
  Vec p_tc, p_tp, p_ic, p_ip;
  Vec fl_tc, fl_tp, fl_ic, fl_ip;
  // fluid properties (hydro2d.hpp:1404-1430)
  Vec rho, rho_s, mu_s, vsrc, msrc, fx, fy, stx, sty;
  Vec pd_tc[HG_MAX_PHASES], pd_tp[HG_MAX_PHASES], pd_ic[HG_MAX_PHASES], pd_ip[HG_MAX_PHASES];
  Vec alpha[HG_MAX_PHASES];
  int fluid_iter = 0;     // FluidSimple iteration count
  int cd_iter = 0;        // ConvectionDiffusionImplicit iteration count
  long pinned = -1;       // fixed-pressure cell
  // systems of the last iteration
  Sys vel_sys;
  Vec vel_k[2];
  Sys p_sys;
  Vec a_p;  // diagonal coefficient
  Timers timers;
  // running part of CalcStat: boundary volume / mass in and out per phase,
  // the output offset meshpos and the value the last statistics used
  double bacc[HG_MAX_PHASES][4] = {};
  // heat: temperature time layers, per-face conditions (boundary faces:
  // kind 0 value fixed, 1 derivative fixed), unsteady coefficients
  Vec t_tc, t_tp;
  std::vector<int> hkind;
  Vec hval;
  std::vector<double> hbdf;
  double meshpos[2] = {0., 0.}, meshpos_stat[2] = {0., 0.};
  // fluid conditions per boundary face: the velocity (walls and inlets
  // fixed, outlets updated every iteration) and, at inlets, the phases'
  // partial densities
  Vec bvel[2];
  Vec pdb[HG_MAX_PHASES];
};

namespace {

SideBc VelBc(const or_ctx* o, int comp) {
  SideBc b;
  for (int s = 0; s < 4; ++s) {
    b.kind[s] = BC_VALUE;
    b.val[s] = o->cfg.bc_velocity[s][comp];
  }
  b.fval = o->bvel[comp].data();
  return b;
}

// Partial density conditions of phase i (hydro2d.hpp:565-583): the inlet
// value on inlet faces, zero derivative elsewhere.
SideBc PdBc(const or_ctx* o, int i) {
  SideBc b = MakeBc(BC_DERIV, 0.);
  for (int s = 0; s < 4; ++s)
    if (o->cfg.bc_kind[s] == HG_BC_INLET) b.kind[s] = BC_VALUE;
  b.fval = o->pdb[i].data();
  return b;
}

// Fluid conditions per face (hydro2d.hpp:370-392, fluid.hpp:394-415): walls
// and inlets hold their velocity, outlets start at zero.
void InitBoundaryVelocity(or_ctx* o) {
  const Mesh& m = o->m;
  for (int d = 0; d < 2; ++d) o->bvel[d].assign(m.nf, 0.);
  for (size_t f = 0; f < m.nf; ++f) {
    if (m.Inner(f)) continue;
    const int s = m.Side(f);
    if (o->cfg.bc_kind[s] != HG_BC_OUTLET)
      for (int d = 0; d < 2; ++d) o->bvel[d][f] = o->cfg.bc_velocity[s][d];
  }
}

// The inlet partial densities: those of the adjacent cells now
// (hydro2d.hpp:565-583).
void InitBoundaryPd(or_ctx* o) {
  const Mesh& m = o->m;
  for (int i = 0; i < HG_MAX_PHASES; ++i) o->pdb[i].assign(m.nf, 0.);
  for (size_t f = 0; f < m.nf; ++f) {
    if (m.Inner(f) || o->cfg.bc_kind[m.Side(f)] != HG_BC_INLET) continue;
    const size_t c = size_t(m.ValidId(f) == 0 ? m.fcm[f] : m.fcp[f]);
    for (int i = 0; i < o->np; ++i) o->pdb[i][f] = o->pd_tc[i][c];
  }
}

// FluidSimple::UpdateOutletBaseConditions (fluid.hpp:542-600), boundary
// faces in face order: outlet faces take the adjacent cell's iter_curr
// velocity; the inflow (inlet faces' u . S toward the domain, plus the
// volume sources) minus the outflow over the outlet area is added along
// each outlet face's normal.
void UpdateOutlets(or_ctx* o) {
  const Mesh& m = o->m;
  bool any = false;
  for (int s = 0; s < 4; ++s) any = any || o->cfg.bc_kind[s] == HG_BC_OUTLET;
  if (!any) return;
  double inflow = 0., outflow = 0., area = 0.;
  for (size_t f = 0; f < m.nf; ++f) {
    if (m.Inner(f)) continue;
    const int kind = o->cfg.bc_kind[m.Side(f)];
    const int id = m.ValidId(f);
    if (kind == HG_BC_OUTLET) {
      const size_t c = size_t(id == 0 ? m.fcm[f] : m.fcp[f]);
      const double factor = (id == 0 ? 1. : -1.);
      o->bvel[0][f] = o->u_ic[0][c];
      o->bvel[1][f] = o->u_ic[1][c];
      outflow += Dot(V2{o->bvel[0][f], o->bvel[1][f]}, m.surf[f]) * factor;
      area += m.area[f];
    } else if (kind == HG_BC_INLET) {
      const double factor = (id == 0 ? -1. : 1.);
      inflow += Dot(V2{o->bvel[0][f], o->bvel[1][f]}, m.surf[f]) * factor;
    }
  }
  for (size_t c = 0; c < m.nc; ++c) inflow += o->vsrc[c] * m.vol[c];
  const double corr = (inflow - outflow) / area;
  for (size_t f = 0; f < m.nf; ++f) {
    if (m.Inner(f) || o->cfg.bc_kind[m.Side(f)] != HG_BC_OUTLET) continue;
    const double factor = (m.ValidId(f) == 0 ? 1. : -1.);
    const V2 nrm{m.surf[f].x / m.area[f], m.surf[f].y / m.area[f]};
    o->bvel[0][f] = o->bvel[0][f] + (nrm.x * corr) * factor;
    o->bvel[1][f] = o->bvel[1][f] + (nrm.y * corr) * factor;
  }
}

double DistV(V2 a, V2 b) { return Norm(V2{b.x - a.x, b.y - a.y}); }

// FluidSimple initial flux (fluid.hpp:770-784): u_f . S - meshvel . S.
void InitialFlux(or_ctx* o) {
  const Mesh& m = o->m;
  Vec ufx(m.nf), ufy(m.nf);
  Interpolate(m, o->u_tc[0].data(), ufx.data(), VelBc(o, 0));
  Interpolate(m, o->u_tc[1].data(), ufy.data(), VelBc(o, 1));
  V2 mv{o->cfg.meshvel[0], o->cfg.meshvel[1]};
  o->fl_tc.assign(m.nf, 0.);
  for (size_t f = 0; f < m.nf; ++f) o->fl_tc[f] = Dot(V2{ufx[f], ufy[f]}, m.surf[f]);
  for (size_t f = 0; f < m.nf; ++f) o->fl_tc[f] -= Dot(mv, m.surf[f]);
  o->fl_tp = o->fl_tc;
}

// hydro::UpdateFluidProperties (hydro2d.hpp:1404-1430) with zero mass sources.
void UpdateProperties(or_ctx* o) {
  const Mesh& m = o->m;
  const hg_config& cfg = o->cfg;
  const size_t nc = m.nc;
  // CalcPhasesVolumeFraction (hydro2d.hpp:980-997)
  for (size_t c = 0; c < nc; ++c) {
    double sum = 0.;
    for (int i = 0; i < o->np; ++i) {
      o->alpha[i][c] = o->pd_tc[i][c] / cfg.density[i];
      sum += o->alpha[i][c];
    }
    for (int i = 0; i < o->np; ++i) o->alpha[i][c] /= sum;
  }
  // GetVolumeAveraged (hydro2d.hpp:1221-1246)
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
  o->msrc.assign(nc, 0.);
  o->vsrc.assign(nc, 0.);
  // CalcForce (hydro2d.hpp:1308-1376)
  o->fx.resize(nc);
  o->fy.resize(nc);
  for (size_t c = 0; c < nc; ++c) {
    o->fx[c] = cfg.gravity[0] * o->rho[c] + cfg.force[0];
    o->fy[c] = cfg.gravity[1] * o->rho[c] + cfg.force[1];
  }
  o->stx.assign(nc, 0.);
  o->sty.assign(nc, 0.);
  if (o->np >= 2) {
    const Vec& a = o->alpha[1];
    Vec af(m.nf), gx(nc), gy(nc), gfx(m.nf), gfy(m.nf);
    Interpolate(m, a.data(), af.data(), MakeBc(BC_DERIV, 0.));
    Gradient(m, af.data(), gx.data(), gy.data());
    Interpolate(m, gx.data(), gfx.data(), MakeBc(BC_DERIV, 0.));
    Interpolate(m, gy.data(), gfy.data(), MakeBc(BC_DERIV, 0.));
    double sigma = cfg.sigma;
    for (size_t c = 0; c < nc; ++c) {
      V2 fsum{0, 0};
      for (int k = 0; k < 4; ++k) {
        size_t f = m.Face(c, k);
        V2 g{gfx[f], gfy[f]};
        V2 nn = g;
        double d = Norm(nn) + 1e-6;
        nn.x /= d;
        nn.y /= d;
        V2 s = m.OutSurf(c, k);
        double sn = Dot(s, nn);
        fsum.x += g.x * sn;
        fsum.y += g.y * sn;
        double gn = Norm(g);
        fsum.x -= s.x * gn;
        fsum.y -= s.y * gn;
      }
      fsum.x /= m.vol[c];
      fsum.y /= m.vol[c];
      o->stx[c] = fsum.x * sigma;
      o->sty[c] = fsum.y * sigma;
    }
  }
  Smooth(m, o->fx, cfg.force_smooth_times);
  Smooth(m, o->fy, cfg.force_smooth_times);
}

}  // namespace

// ---------------------------------------------------------------------------

extern "C" {

// The running part of hydro::CalcStat (hydro2d.hpp:1474-1528), after every
// property update: per phase, the boundary faces in face order (inner faces
// skipped), the time_curr flux toward the valid neighbour cell (factor +1
// when it is cell 1, -1 when cell 0) times that cell's volume fraction,
// summed into "in" when >= 0 and subtracted into "out" otherwise; then the
// totals += flux * dt.  Last, meshpos += meshvel * dt (meshvel_output).
static void RunningStats(or_ctx* o) {
  const Mesh& m = o->m;
  for (int i = 0; i < o->np; ++i) {
    double vin = 0., vout = 0.;
    for (size_t f = 0; f < m.nf; ++f) {
      if (m.Inner(f)) continue;
      const int id = m.ValidId(f);
      const size_t c = size_t(id == 1 ? m.fcp[f] : m.fcm[f]);
      const double factor = id == 1 ? 1. : -1.;
      const double to_cell = o->fl_tc[f] * factor;
      const double vf = o->alpha[i][c];
      if (to_cell >= 0) vin += to_cell * vf;
      else vout -= to_cell * vf;
    }
    const double min = vin * o->cfg.density[i], mout = vout * o->cfg.density[i];
    o->bacc[i][0] += vin * o->dt;
    o->bacc[i][1] += vout * o->dt;
    o->bacc[i][2] += min * o->dt;
    o->bacc[i][3] += mout * o->dt;
  }
  for (int d = 0; d < 2; ++d) o->meshpos_stat[d] = o->meshpos[d];
  if (o->cfg.meshvel_output)
    for (int d = 0; d < 2; ++d) o->meshpos[d] += o->cfg.meshvel[d] * o->dt;
}

static void InitHeat(or_ctx* o);

int or_create(const hg_config* cfg_in, or_ctx** out, char* err, size_t errlen) {
  auto fail = [&](const char* msg) {
    if (err && errlen) std::snprintf(err, errlen, "%s", msg);
    return int(HG_ERR_UNSUPPORTED);
  };
  const hg_config& cfg = *cfg_in;
  if (cfg.nx < 2 || cfg.ny < 2) return fail("Nx, Ny must be >= 2");
  if (cfg.num_phases < 1 || cfg.num_phases > HG_MAX_PHASES) return fail("num_phases 1..3");
  if (cfg.force_geometric_average) return fail("force_geometric_average=1 unsupported");
  or_ctx* o = new or_ctx();
  o->cfg = cfg;
  o->np = cfg.num_phases;
  Mesh& m = o->m;
  m.Init(cfg.nx, cfg.ny, cfg.a, cfg.b);
  o->dt = cfg.dt;
  o->dt_adv = cfg.dt * cfg.advection_dt_factor;
  o->bdf = DerivCoeffs(0., {-2. * o->dt, -o->dt, 0.}, cfg.time_second_order ? 0 : 1);
  const size_t nc = m.nc;
  if (cfg.heat_enable) InitHeat(o);
  // Initial velocity (hydro2d.hpp:316-367)
  for (int d = 0; d < 2; ++d) o->u_tc[d].assign(nc, cfg.initial_velocity[d]);
  if (cfg.has_initial_sin) {
    double pi = std::atan(1.) * 4.;
    double kx = cfg.initial_sin_n[0] * (2. * pi / cfg.initial_sin_lambda);
    double ky = cfg.initial_sin_n[1] * (2. * pi / cfg.initial_sin_lambda);
    for (size_t c = 0; c < nc; ++c) {
      double sv = std::sin(Dot(V2{kx, ky}, m.cc[c]) - cfg.initial_sin_phase);
      o->u_tc[0][c] *= sv;
      o->u_tc[1][c] *= sv;
    }
  }
  for (int d = 0; d < 2; ++d) o->u_tp[d] = o->u_tc[d];
  // Fixed pressure point: FindNearestCell (mesh.hpp:413-422)
  if (cfg.has_pressure_fixed_point) {
    V2 pt{cfg.pressure_fixed_point[0], cfg.pressure_fixed_point[1]};
    size_t best = 0;
    for (size_t c = 0; c < nc; ++c)
      if (DistV(pt, m.cc[c]) < DistV(pt, m.cc[best])) best = c;
    o->pinned = long(best);
  }
  o->p_tc.assign(nc, 0.);
  o->p_tp = o->p_tc;
  InitBoundaryVelocity(o);
  InitialFlux(o);
  // Initial partial densities (hydro2d.hpp:494-566)
  for (int i = 0; i < o->np; ++i)
    o->pd_tc[i].assign(nc, cfg.density[i] * cfg.initial_volume_fraction[i]);
  for (size_t c = 0; c < nc; ++c) {
    V2 x = m.cc[c];
    auto inside = [&](const double* lo, const double* hi) {
      return !(x.x < lo[0] || hi[0] < x.x) && !(x.y < lo[1] || hi[1] < x.y);
    };
    if (inside(cfg.a2, cfg.b2)) {
      if (o->np < 3) {
        delete o;
        return fail("block A2-B2 holds cells but num_phases < 3");
      }
      o->pd_tc[2][c] = cfg.density[2];
    } else if (o->np >= 2 && inside(cfg.a1, cfg.b1)) {
      o->pd_tc[1][c] = cfg.density[1];
    } else if (o->np >= 2 && DistV(x, V2{cfg.ic[0], cfg.ic[1]}) < cfg.ir) {
      o->pd_tc[1][c] = cfg.density[1];
    } else if (o->np >= 2 && DistV(x, V2{cfg.ic2[0], cfg.ic2[1]}) < cfg.ir2) {
      o->pd_tc[1][c] = cfg.density[1];
    }
  }
  for (int i = 1; i < o->np; ++i) Smooth(m, o->pd_tc[i], cfg.initial_volume_fraction_smooth_times);
  for (size_t c = 0; c < nc; ++c) {
    double vs = 0.;
    for (int i = 1; i < o->np; ++i) vs += o->pd_tc[i][c] / cfg.density[i];
    o->pd_tc[0][c] = (1. - vs) * cfg.density[0];
  }
  for (int i = 0; i < o->np; ++i) {
    o->pd_tp[i] = o->pd_tc[i];
    o->alpha[i].assign(nc, 0.);
  }
  InitBoundaryPd(o);
  UpdateProperties(o);
  RunningStats(o);
  *out = o;
  return HG_OK;
}

void or_destroy(or_ctx* o) { delete o; }
int or_num_cells(const or_ctx* o) { return int(o->m.nc); }
int or_num_faces(const or_ctx* o) { return int(o->m.nf); }

static Vec* FieldPtr(or_ctx* o, int field, bool write) {
  if (field == HG_FIELD_VELOCITY_X) return &o->u_tc[0];
  if (field == HG_FIELD_VELOCITY_Y) return &o->u_tc[1];
  if (field == HG_FIELD_PRESSURE) return &o->p_tc;
  if (field == HG_FIELD_VOLUME_FLUX) return &o->fl_tc;
  if (field == HG_FIELD_TEMPERATURE && o->cfg.heat_enable) return &o->t_tc;
  if (field >= HG_FIELD_PARTIAL_DENSITY_0 && field < HG_FIELD_PARTIAL_DENSITY_0 + o->np)
    return &o->pd_tc[field - HG_FIELD_PARTIAL_DENSITY_0];
  if (write) return nullptr;
  if (field >= HG_FIELD_VOLUME_FRACTION_0 && field < HG_FIELD_VOLUME_FRACTION_0 + o->np)
    return &o->alpha[field - HG_FIELD_VOLUME_FRACTION_0];
  if (field == HG_FIELD_DENSITY) return &o->rho;
  if (field == HG_FIELD_DENSITY_SMOOTH) return &o->rho_s;
  if (field == HG_FIELD_VISCOSITY_SMOOTH) return &o->mu_s;
  if (field == HG_FIELD_FORCE_X) return &o->fx;
  if (field == HG_FIELD_FORCE_Y) return &o->fy;
  if (field == HG_FIELD_VOLUME_FLUX_ITER) return &o->fl_ic;
  return nullptr;
}

int or_set_field(or_ctx* o, int field, const double* host, size_t n) {
  Vec* v = FieldPtr(o, field, true);
  if (!v || v->size() != n) return HG_ERR_BAD_ARG;
  std::copy(host, host + n, v->begin());
  if (field == HG_FIELD_VELOCITY_X) o->u_tp[0] = *v;
  if (field == HG_FIELD_VELOCITY_Y) o->u_tp[1] = *v;
  if (field == HG_FIELD_PRESSURE) o->p_tp = *v;
  if (field == HG_FIELD_VOLUME_FLUX) o->fl_tp = *v;
  if (field == HG_FIELD_TEMPERATURE) o->t_tp = *v;
  if (field >= HG_FIELD_PARTIAL_DENSITY_0 && field < HG_FIELD_PARTIAL_DENSITY_0 + o->np)
    o->pd_tp[field - HG_FIELD_PARTIAL_DENSITY_0] = *v;
  return HG_OK;
}

int or_get_field(or_ctx* o, int field, double* host, size_t n) {
  Vec* v = FieldPtr(o, field, false);
  if (!v || v->size() != n) return HG_ERR_BAD_ARG;
  std::copy(v->begin(), v->end(), host);
  return HG_OK;
}

int or_refresh(or_ctx* o) {
  InitBoundaryVelocity(o);
  InitBoundaryPd(o);
  InitialFlux(o);
  UpdateProperties(o);
  std::memset(o->bacc, 0, sizeof(o->bacc));
  o->meshpos[0] = o->meshpos[1] = 0.;
  RunningStats(o);
  return HG_OK;
}

// FluidSimple::StartStep (fluid.hpp:793-812) with the component solvers'
// StartStep (conv_diff.hpp:117-128).
int or_fluid_start_step(or_ctx* o) {
  const hg_config& cfg = o->cfg;
  o->fluid_iter = 0;
  for (double v : o->p_tc)
    if (!(v * 0. == 0.)) return HG_ERR_NAN;
  for (int d = 0; d < 2; ++d) {
    for (double v : o->u_tc[d])
      if (!(v * 0. == 0.)) return HG_ERR_NAN;
    o->u_ic[d] = o->u_tc[d];
    for (size_t c = 0; c < o->m.nc; ++c)
      o->u_ic[d][c] += (o->u_tc[d][c] - o->u_tp[d][c]) * cfg.guess_extrapolation;
  }
  o->cd_iter = 0;
  o->p_ic = o->p_tc;
  o->fl_ic = o->fl_tc;
  for (size_t c = 0; c < o->m.nc; ++c)
    o->p_ic[c] += (o->p_tc[c] - o->p_tp[c]) * cfg.guess_extrapolation;
  for (size_t f = 0; f < o->m.nf; ++f)
    o->fl_ic[f] += (o->fl_tc[f] - o->fl_tp[f]) * cfg.guess_extrapolation;
  return HG_OK;
}

static double ConvergenceIndicator(const or_ctx* o) {
  if (o->cd_iter == 0) return 1.;
  double r = 0.;
  for (size_t c = 0; c < o->m.nc; ++c) {
    V2 a{o->u_ic[0][c], o->u_ic[1][c]}, b{o->u_ip[0][c], o->u_ip[1][c]};
    r = std::max(r, DistV(a, b));
  }
  return r;
}

int or_fluid_is_converged(const or_ctx* o) {
  return (o->fluid_iter >= o->cfg.num_iterations_limit ||
          ConvergenceIndicator(o) < o->cfg.convergence_tolerance)
             ? 1
             : 0;
}

// Solves rows with the solver named by `which` (hg_solver).
static SolveResult RunSolver(const or_ctx* o, int which, const Sys& a, const Vec& k, Vec& x,
                             bool singular) {
  const hg_config& cfg = o->cfg;
  SolveResult r;
  switch (which) {
    case HG_SOLVER_LU: SolveLu(a, k, x); break;
    case HG_SOLVER_LU_RB: SolveLuRb(a, k, x); break;
    case HG_SOLVER_GAUSS_SEIDEL:
      r = SolveGs(a, k, x, cfg.lu_relaxed_tolerance, cfg.lu_relaxed_num_iters_limit,
                  cfg.lu_relaxed_relaxation_factor);
      break;
    case HG_SOLVER_JACOBI:
      r = SolveJacobi(a, k, x, cfg.lu_relaxed_tolerance, cfg.lu_relaxed_num_iters_limit,
                      cfg.lu_relaxed_relaxation_factor);
      break;
    case HG_SOLVER_LU_RELAXED:
      r = SolveLuRelaxed(a, k, x, cfg.lu_relaxed_tolerance, cfg.lu_relaxed_num_iters_limit,
                         cfg.lu_relaxed_relaxation_factor);
      break;
    default:
      r = SolveCg(a, k, x, cfg.pressure_tolerance_gpu, cfg.pressure_max_iters_gpu, singular);
      break;
  }
  return r;
}

// FluidSimple::MakeIteration (fluid.hpp:814-1158).
int or_fluid_iterate(or_ctx* o, hg_iter_stats* stats) {
  const hg_config& cfg = o->cfg;
  const Mesh& m = o->m;
  const size_t nc = m.nc, nf = m.nf;
  const long lnc = long(nc), lnf = long(nf);
  o->p_ip = o->p_ic;
  o->fl_ip = o->fl_ic;
  UpdateOutlets(o);

  Vec ffex(nf), ffey(nf), frx(nc), fry(nc), ffrx(nf), ffry(nf), ffmu(nf);
  {
    ScopedTimer t(&o->timers, "fluid.1.force-correction");
    // CalcExtForce (fluid.hpp:602-631)
    SideBc d0 = MakeBc(BC_DERIV, 0.);
    Interpolate(m, o->fx.data(), ffex.data(), d0);
    Interpolate(m, o->fy.data(), ffey.data(), d0);
#pragma omp parallel for
    for (long c = 0; c < lnc; ++c) {
      V2 sum{0, 0};
      for (int k = 0; k < 4; ++k) {
        size_t f = m.Face(c, k);
        double dn = Dot(V2{ffex[f], ffey[f]}, m.normal[f]);
        double dist = DistV(m.cc[c], m.fc[f]);
        sum.x += m.surf[f].x * dn * dist;
        sum.y += m.surf[f].y * dn * dist;
      }
      frx[c] = sum.x / m.vol[c];
      fry[c] = sum.y / m.vol[c];
    }
    Interpolate(m, frx.data(), ffrx.data(), d0);
    Interpolate(m, fry.data(), ffry.data(), d0);
  }
  // CalcKinematicViscosity (fluid.hpp:632-641)
  Interpolate(m, o->mu_s.data(), ffmu.data(), MakeBc(BC_DERIV, 0.));

  Vec ffp(nf), gpx(nc), gpy(nc), ffgpx(nf), ffgpy(nf);
  {
    ScopedTimer t(&o->timers, "fluid.0.pressure-gradient");
    Interpolate(m, o->p_ip.data(), ffp.data(), MakeBc(BC_EXTRAP));
    Gradient(m, ffp.data(), gpx.data(), gpy.data());
    Interpolate(m, gpx.data(), ffgpx.data(), MakeBc(BC_DERIV, 0.));
    Interpolate(m, gpy.data(), ffgpy.data(), MakeBc(BC_DERIV, 0.));
  }

  // Explicit viscosity (fluid.hpp:835-854)
  Vec fox(nc, 0.), foy(nc, 0.);
  {
    ScopedTimer t(&o->timers, "fluid.1a.explicit-viscosity");
    Vec ff(nf), gx(nc), gy(nc), gfx(nf), gfy(nf);
    for (int d = 0; d < 2; ++d) {
      Interpolate(m, o->u_ic[d].data(), ff.data(), VelBc(o, d));
      Gradient(m, ff.data(), gx.data(), gy.data());
      Interpolate(m, gx.data(), gfx.data(), MakeBc(BC_DERIV, 0.));
      Interpolate(m, gy.data(), gfy.data(), MakeBc(BC_DERIV, 0.));
#pragma omp parallel for
      for (long c = 0; c < lnc; ++c) {
        V2 sum{0, 0};
        for (int k = 0; k < 4; ++k) {
          size_t f = m.Face(c, k);
          V2 os = m.OutSurf(c, k);
          double w = ffmu[f] * (d == 0 ? os.x : os.y);
          sum.x += gfx[f] * w;
          sum.y += gfy[f] * w;
        }
        fox[c] += sum.x / m.vol[c];
        foy[c] += sum.y / m.vol[c];
      }
    }
  }
  // Force assembly (fluid.hpp:857-870)
#pragma omp parallel for
  for (long c = 0; c < lnc; ++c) {
    double s = o->rho_s[c] * o->vsrc[c] - o->msrc[c];
    fox[c] += ((gpx[c] * (-1.) + frx[c]) + o->stx[c]) + o->u_ic[0][c] * s;
    foy[c] += ((gpy[c] * (-1.) + fry[c]) + o->sty[c]) + o->u_ic[1][c] * s;
  }

  // Momentum: ConvectionDiffusionScalarImplicit::MakeIteration per component
  // (conv_diff.hpp:130-251).  Coefficients are the same for both
  // components; only the constants differ.
  Sys& A = o->vel_sys;
  A.Resize(m.nx, m.ny);
  o->vel_k[0].assign(nc, 0.);
  o->vel_k[1].assign(nc, 0.);
  int vel_iters[2] = {1, 1};
  {
    ScopedTimer t(&o->timers, "fluid.2.convection-diffusion");
    const Vec& probe = o->fl_ip;
    for (int d = 0; d < 2; ++d) {
      o->u_ip[d] = o->u_ic[d];
      const Vec& up = o->u_ip[d];
      Vec ff(nf), gx(nc), gy(nc);
      SideBc vbc = VelBc(o, d);
      Interpolate(m, up.data(), ff.data(), vbc);
      Gradient(m, ff.data(), gx.data(), gy.data());
      // face expressions: convective (am, ap, ak), diffusive (dm, dp, dk)
      Vec cam(nf), cap(nf), cak(nf), dcm(nf), dcp(nf), dck(nf);
#pragma omp parallel for
      for (long f = 0; f < lnf; ++f) {
        double F = probe[f];
        double mu = ffmu[f];
        double A_ = m.area[f];
        if (m.Inner(f)) {
          long cm = m.fcm[f], cp = m.fcp[f];
          double am, ap, ak;
          if (F > 1e-10) {
            am = 1.;
            ap = 0.;
            ak = -Dot(V2{gx[cm], gy[cm]}, m.to0[f]);
          } else if (F < -1e-10) {
            am = 0.;
            ap = 1.;
            ak = -Dot(V2{gx[cp], gy[cp]}, m.to1[f]);
          } else {
            am = 0.5;
            ap = 0.5;
            ak = 0.;
          }
          cam[f] = am * F;
          cap[f] = ap * F;
          cak[f] = ak * F;
          V2 dd{m.to1[f].x - m.to0[f].x, m.to1[f].y - m.to0[f].y};
          double al = 1. / Norm(dd);
          dcm[f] = ((-al) * (-mu)) * A_;
          dcp[f] = (al * (-mu)) * A_;
          dck[f] = (0. * (-mu)) * A_;
        } else {
          int id = m.ValidId(f);
          double val = vbc.fval ? vbc.fval[f] : vbc.val[m.Side(f)];
          cam[f] = cap[f] = 0.;
          cak[f] = val * F;
          double factor = (id == 0 ? 1. : -1.);
          double al = 1. / Norm(m.ToCell(f, id)) * factor;
          // boundary term sits on the valid neighbour
          double dcc = ((-al) * (-mu)) * A_;
          if (id == 0) {
            dcm[f] = dcc;
            dcp[f] = 0.;
          } else {
            dcm[f] = 0.;
            dcp[f] = dcc;
          }
          dck[f] = ((al * val) * (-mu)) * A_;
        }
      }
      const std::vector<double>& co = o->bdf;
      const double relax = cfg.velocity_relaxation_factor;
      Vec& kk = o->vel_k[d];
#pragma omp parallel for
      for (long c = 0; c < lnc; ++c) {
        // merged sums over faces W, E, S, N (conv_diff.hpp:187-200)
        double Cc = 0., Ck = 0., Dc = 0., Dk = 0.;
        double Cnb[4] = {0, 0, 0, 0}, Dnb[4] = {0, 0, 0, 0};
        for (int k = 0; k < 4; ++k) {
          size_t f = m.Face(c, k);
          double o_ = Mesh::Out(k);
          bool self_is_m = (k == 1 || k == 3);  // E, N faces: c is cm
          double cself = self_is_m ? cam[f] : cap[f];
          double cnb = self_is_m ? cap[f] : cam[f];
          double dself = self_is_m ? dcm[f] : dcp[f];
          double dnb = self_is_m ? dcp[f] : dcm[f];
          if (m.Inner(f)) {
            Cc += cself * o_;
            Cnb[k] = cnb * o_;
            Dnb[k] = dnb * o_;
          }
          Dc += dself * o_;
          Ck += cak[f] * o_;
          Dk += dck[f] * o_;
        }
        double V = m.vol[c], rho = o->rho_s[c];
        // unsteady + cflux/V, then *rho, then dflux/V + that, minus source
        double cen = (Dc / V) + ((Cc / V + co[2]) * rho);
        double nb[4];
        for (int k = 0; k < 4; ++k) nb[k] = (Dnb[k] / V) + ((Cnb[k] / V) * rho);
        double unst = co[0] * (d == 0 ? o->u_tp[0][c] : o->u_tp[1][c]) +
                      co[1] * (d == 0 ? o->u_tc[0][c] : o->u_tc[1][c]);
        double src = d == 0 ? fox[c] : foy[c];
        double kon = ((Dk / V) + ((Ck / V + unst) * rho)) + (-src);
        // delta form: constant = row evaluated at u_prev (order S, W, C, E, N)
        int i = int(c % m.nx), j = int(c / m.nx);
        double ev = kon;
        if (j > 0) ev += up[c - m.nx] * nb[2];
        if (i > 0) ev += up[c - 1] * nb[0];
        ev += up[c] * cen;
        if (i < m.nx - 1) ev += up[c + 1] * nb[1];
        if (j < m.ny - 1) ev += up[c + m.nx] * nb[3];
        kk[c] = ev;
        if (d == 0) {
          A.w[c] = i > 0 ? nb[0] : 0.;
          A.e[c] = i < m.nx - 1 ? nb[1] : 0.;
          A.s[c] = j > 0 ? nb[2] : 0.;
          A.n[c] = j < m.ny - 1 ? nb[3] : 0.;
          A.c[c] = cen / relax;
        }
      }
    }
    for (int d = 0; d < 2; ++d) {
      Vec corr;
      SolveResult r = RunSolver(o, cfg.linear_solver_velocity, A, o->vel_k[d], corr, false);
      vel_iters[d] = r.iters;
      for (size_t c = 0; c < nc; ++c) o->u_ic[d][c] = o->u_ip[d][c] + corr[c];
    }
    o->cd_iter += 1;
  }

  // Diagonal coefficient (fluid.hpp:876-887)
  o->a_p.assign(nc, 0.);
  for (size_t c = 0; c < nc; ++c) {
    double cs = A.CoeffSum(c);
    double sum = 0.;
    sum += cs;
    sum += cs;
    o->a_p[c] = sum / 2.;
  }
  Vec ffd(nf);
  Interpolate(m, o->a_p.data(), ffd.data(), MakeBc(BC_NONE));
  Vec fux(nf), fuy(nf), fstar(nf), fcoef(nf, 0.);
  {
    ScopedTimer t(&o->timers, "fluid.3.momentum-interpolation");
    Interpolate(m, o->u_ic[0].data(), fux.data(), VelBc(o, 0));
    Interpolate(m, o->u_ic[1].data(), fuy.data(), VelBc(o, 1));
    const double rc = cfg.rhie_chow_factor;
    V2 mv{cfg.meshvel[0], cfg.meshvel[1]};
#pragma omp parallel for
    for (long f = 0; f < lnf; ++f) {
      double vfi = Dot(V2{fux[f], fuy[f]}, m.surf[f]);
      if (m.Inner(f)) {
        long cm = m.fcm[f], cp = m.fcp[f];
        V2 dd{m.to1[f].x - m.to0[f].x, m.to1[f].y - m.to0[f].y};
        double wide = Dot(V2{ffgpx[f] - ffrx[f], ffgpy[f] - ffry[f]}, m.surf[f]);
        double compact = (o->p_ip[cp] - o->p_ip[cm]) / Norm(dd) * m.area[f] -
                         Dot(V2{ffex[f], ffey[f]}, m.surf[f]);
        fstar[f] = vfi + rc * (wide - compact) / ffd[f] + 0;
      } else {
        fstar[f] = vfi;
      }
    }
    for (size_t f = 0; f < nf; ++f) fstar[f] -= Dot(mv, m.surf[f]);
  }
  {
    ScopedTimer t(&o->timers, "fluid.4.volume-flux");
#pragma omp parallel for
    for (long f = 0; f < lnf; ++f) {
      if (m.Inner(f)) {
        V2 dd{m.to1[f].x - m.to0[f].x, m.to1[f].y - m.to0[f].y};
        fcoef[f] = -m.area[f] / (Norm(dd) * ffd[f]);
      }
    }
  }
  // Pressure-correction system (fluid.hpp:972-1027)
  Sys& P = o->p_sys;
  P.Resize(m.nx, m.ny);
  {
    ScopedTimer t(&o->timers, "fluid.5.pressure-system");
#pragma omp parallel for
    for (long c = 0; c < lnc; ++c) {
      double cen = 0., kon = 0.;
      double nb[4] = {0, 0, 0, 0};
      for (int k = 0; k < 4; ++k) {
        size_t f = m.Face(c, k);
        double o_ = Mesh::Out(k);
        bool self_is_m = (k == 1 || k == 3);
        if (m.Inner(f)) {
          double cself = self_is_m ? -fcoef[f] : fcoef[f];
          double cnb = self_is_m ? fcoef[f] : -fcoef[f];
          cen += cself * o_;
          nb[k] = cnb * o_;
        }
        kon += fstar[f] * o_;
      }
      kon += -(o->vsrc[c] * m.vol[c]);
      P.w[c] = nb[0];
      P.e[c] = nb[1];
      P.s[c] = nb[2];
      P.n[c] = nb[3];
      P.c[c] = cen;
      P.k[c] = kon;
    }
    if (o->pinned >= 0) {
      // Expression::SetKnownValue (linear.hpp:237-249) on every other row.
      size_t pc = size_t(o->pinned);
      double val = cfg.pressure_fixed_value;
      for (size_t c = 0; c < nc; ++c) {
        if (c == pc) continue;
        if (c + m.nx == pc) { P.k[c] += val * P.n[c]; P.n[c] = 0.; }
        if (c + 1 == pc && c % m.nx != size_t(m.nx - 1)) { P.k[c] += val * P.e[c]; P.e[c] = 0.; }
        if (c == pc + 1 && c % m.nx != 0) { P.k[c] += val * P.w[c]; P.w[c] = 0.; }
        if (c == pc + m.nx) { P.k[c] += val * P.s[c]; P.s[c] = 0.; }
      }
      P.s[pc] = P.w[pc] = P.e[pc] = P.n[pc] = 0.;
      P.c[pc] = 1.;
      P.k[pc] = -val;
    }
  }
  Vec pc;
  SolveResult pr;
  {
    ScopedTimer t(&o->timers, "fluid.6.pressure-solve");
    pr = RunSolver(o, cfg.linear_solver_pressure, P, P.k, pc, o->pinned < 0);
  }
  {
    ScopedTimer t(&o->timers, "fluid.7.correction");
    for (size_t c = 0; c < nc; ++c)
      o->p_ic[c] = o->p_ip[c] + cfg.pressure_relaxation_factor * pc[c];
    Vec ffpc(nf), gx(nc), gy(nc);
    Interpolate(m, pc.data(), ffpc.data(), MakeBc(BC_EXTRAP));
    Gradient(m, ffpc.data(), gx.data(), gy.data());
    for (size_t c = 0; c < nc; ++c) {
      o->u_ic[0][c] += gx[c] / (-o->a_p[c]);
      o->u_ic[1][c] += gy[c] / (-o->a_p[c]);
    }
    for (size_t f = 0; f < nf; ++f) {
      double v = fstar[f];
      if (m.Inner(f)) {
        v += pc[m.fcm[f]] * (-fcoef[f]);
        v += pc[m.fcp[f]] * fcoef[f];
      }
      o->fl_ic[f] = v;
    }
  }
  if (cfg.simpler) {
    // SIMPLER pressure correction (fluid.hpp:1059-1151): the momentum rows
    // of this iteration evaluated on the velocity change, plus restored
    // force minus the pressure gradient, give face values whose Rhie-Chow
    // balance with the corrected flux is the right-hand side of a second
    // pressure solve on the same rows; its result is added to p.
    ScopedTimer t(&o->timers, "fluid.8.simpler");
    const double rc = cfg.rhie_chow_factor;
    Vec fux(nf), fuy(nf);
    Interpolate(m, o->u_ic[0].data(), fux.data(), VelBc(o, 0));
    Interpolate(m, o->u_ic[1].data(), fuy.data(), VelBc(o, 1));
    Vec dl[2], ev[2];
    for (int d = 0; d < 2; ++d) {
      dl[d].resize(nc);
      for (size_t c = 0; c < nc; ++c) dl[d][c] = o->u_ic[d][c] - o->u_ip[d][c];
    }
    for (int d = 0; d < 2; ++d) {
      ev[d].resize(nc);
      const Vec& x = dl[d];
      for (size_t c = 0; c < nc; ++c) {
        // Expression::Evaluate (linear.hpp:129-135), terms S, W, C, E, N
        const int i = int(c % m.nx), j = int(c / m.nx);
        double r = o->vel_k[d][c];
        if (j > 0) r += x[c - m.nx] * A.s[c];
        if (i > 0) r += x[c - 1] * A.w[c];
        r += x[c] * A.c[c];
        if (i < m.nx - 1) r += x[c + 1] * A.e[c];
        if (j < m.ny - 1) r += x[c + m.nx] * A.n[c];
        ev[d][c] = r;
      }
    }
    for (size_t c = 0; c < nc; ++c) {
      ev[0][c] += frx[c] - gpx[c];
      ev[1][c] += fry[c] - gpy[c];
    }
    Vec fe0(nf), fe1(nf), rhs(nf, 0.);
    Interpolate(m, ev[0].data(), fe0.data(), MakeBc(BC_NONE));
    Interpolate(m, ev[1].data(), fe1.data(), MakeBc(BC_NONE));
    for (size_t f = 0; f < nf; ++f) {
      if (!m.Inner(f)) continue;
      const V2 e{fe0[f] - ffex[f], fe1[f] - ffey[f]};
      rhs[f] = Dot(e, m.surf[f]) / ffd[f] +
               (o->fl_ic[f] - Dot(V2{fux[f], fuy[f]}, m.surf[f])) / rc;
    }
    const long pin = o->pinned;
    for (size_t c = 0; c < nc; ++c) {
      double sum = 0.;
      for (int q = 0; q < 4; ++q) sum += rhs[m.Face(c, q)] * Mesh::Out(q);
      P.k[c] = -sum;
    }
    if (pin >= 0) P.k[size_t(pin)] = -cfg.pressure_fixed_value;
    // constant = the row evaluated at the current pressure (terms removed
    // for a fixed cell stay removed)
    for (size_t c = 0; c < nc; ++c) {
      const int i = int(c % m.nx), j = int(c / m.nx);
      const long lc = long(c);
      double r = P.k[c];
      if (lc == pin) {
        r += o->p_ic[c] * P.c[c];
      } else {
        if (j > 0 && lc - m.nx != pin) r += o->p_ic[c - m.nx] * P.s[c];
        if (i > 0 && lc - 1 != pin) r += o->p_ic[c - 1] * P.w[c];
        r += o->p_ic[c] * P.c[c];
        if (i < m.nx - 1 && lc + 1 != pin) r += o->p_ic[c + 1] * P.e[c];
        if (j < m.ny - 1 && lc + m.nx != pin) r += o->p_ic[c + m.nx] * P.n[c];
      }
      P.k[c] = r;
    }
    Vec pc2;
    SolveResult pr2 = RunSolver(o, cfg.linear_solver_pressure, P, P.k, pc2, pin < 0);
    for (size_t c = 0; c < nc; ++c) o->p_ic[c] += pc2[c];
    pr.iters += pr2.iters;
  }
  o->fluid_iter += 1;
  if (stats) {
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

// FluidSimple::FinishStep (fluid.hpp:1159-1169).
int or_fluid_finish_step(or_ctx* o) {
  o->p_tp = o->p_tc;
  o->fl_tp = o->fl_tc;
  o->p_tc = o->p_ic;
  o->fl_tc = o->fl_ic;
  for (double v : o->p_tc)
    if (!(v * 0. == 0.)) return HG_ERR_NAN;
  for (int d = 0; d < 2; ++d) {
    o->u_tp[d] = o->u_tc[d];
    o->u_tc[d] = o->u_ic[d];
    for (double v : o->u_tc[d])
      if (!(v * 0. == 0.)) return HG_ERR_NAN;
  }
  o->t_fluid += o->dt;
  return HG_OK;
}

// AdvectionSolverMultiExplicit StartStep + MakeIteration + FinishStep
// (advection.hpp:417-545).
int or_advect_substep(or_ctx* o) {
  const hg_config& cfg = o->cfg;
  const Mesh& m = o->m;
  const size_t nc = m.nc, nf = m.nf;
  const long lnc = long(nc), lnf = long(nf);
  for (int i = 0; i < o->np; ++i) {
    o->pd_tp[i] = o->pd_tc[i];
    o->pd_ic[i] = o->pd_tp[i];
  }
  const Vec& flux = o->fl_ic;  // mixture flux (iter_curr layer), slip = 0
  for (int i = 0; i < o->np; ++i) {
    const SideBc d0 = PdBc(o, i);
    o->pd_ip[i] = o->pd_ic[i];
    Vec& curr = o->pd_ic[i];
    Vec vf(nf);
    for (size_t f = 0; f < nf; ++f) vf[f] = flux[f] + 0.;
    int stages = cfg.tvd_split ? 2 : 1;
    for (int stage = 0; stage < stages; ++stage) {
      Vec uf(nf), gx(nc), gy(nc), fu(nf);
      Interpolate(m, curr.data(), uf.data(), d0);
      Gradient(m, uf.data(), gx.data(), gy.data());
      // InterpolateSuperbee (solver.hpp:560-619)
#pragma omp parallel for
      for (long f = 0; f < lnf; ++f) {
        if (m.Inner(f)) {
          long P = m.fcm[f], E = m.fcp[f];
          const Vec& u = curr;
          if (vf[f] > 1e-8) {
            fu[f] = u[P] + 0.5 * Superbee(u[E] - u[P],
                                          -4. * Dot(V2{gx[P], gy[P]}, m.to0[f]) - (u[E] - u[P]));
          } else if (vf[f] < -1e-8) {
            fu[f] = u[E] - 0.5 * Superbee(u[E] - u[P],
                                          4. * Dot(V2{gx[E], gy[E]}, m.to1[f]) - (u[E] - u[P]));
          } else {
            fu[f] = 0.5 * (u[P] + u[E]);
          }
        } else if (d0.kind[m.Side(f)] == BC_VALUE) {
          fu[f] = o->pdb[i][f];  // inlet
        } else {
          int id = m.ValidId(f);
          long cc = id == 0 ? m.fcm[f] : m.fcp[f];
          double factor = (id == 0 ? 1. : -1.);
          double al = Norm(m.ToCell(f, id)) * factor;
          fu[f] = curr[cc] + 0. * al;
        }
      }
      const double dt = o->dt_adv;
#pragma omp parallel for
      for (long c = 0; c < lnc; ++c) {
        double fs = 0.;
        for (int k = 0; k < 4; ++k) {
          if ((k / 2) % stages != stage) continue;
          size_t f = m.Face(c, k);
          fs += fu[f] * vf[f] * Mesh::Out(k);
        }
        curr[c] += -dt / m.vol[c] * fs;
      }
    }
    // Interface sharpening (advection.hpp:482-530)
    Vec ff(nf, 0.);
    if (std::abs(cfg.sharp) > 1e-10) {
      Vec af(nf), gx(nc), gy(nc), gfx(nf), gfy(nf);
      Interpolate(m, curr.data(), af.data(), d0);  // the phase's conditions
      Gradient(m, af.data(), gx.data(), gy.data());
      Interpolate(m, gx.data(), gfx.data(), MakeBc(BC_DERIV, 0.));  // zero derivative
      Interpolate(m, gy.data(), gfy.data(), MakeBc(BC_DERIV, 0.));
      double am = cfg.density[i];
      for (size_t f = 0; f < nf; ++f) {
        V2 n{gfx[f], gfy[f]};
        if (Norm(n) < 1.) continue;
        double d = Norm(n) + 1e-6;
        n.x /= d;
        n.y /= d;
        double nfv = Dot(n, m.normal[f]);
        double uf = vf[f];
        double epsh = cfg.sharp * m.area[f];
        ff[f] = std::abs(uf * nfv) * nfv *
                (epsh * Norm(V2{gfx[f], gfy[f]}) - af[f] * (1. - af[f] / am));
      }
    }
    const double dt = o->dt_adv;
    for (size_t c = 0; c < nc; ++c) {
      curr[c] += dt * o->msrc[c] * 0.;  // phase mass sources are zero on this path
      double sh = 0.;
      for (int k = 0; k < 4; ++k) sh += Mesh::Out(k) * ff[m.Face(c, k)];
      curr[c] += dt * sh / m.vol[c];
    }
  }
  for (int i = 0; i < o->np; ++i) o->pd_tc[i] = o->pd_ic[i];
  o->t_adv += o->dt_adv;
  return HG_OK;
}

int or_update_properties(or_ctx* o) {
  UpdateProperties(o);
  RunningStats(o);
  return HG_OK;
}

// hydro::CalcStat (hydro2d.hpp:1432-1529).
int or_calc_stats(or_ctx* o, hg_stats* st) {
  const Mesh& m = o->m;
  std::memset(st, 0, sizeof(*st));
  for (int i = 0; i < o->np; ++i) {
    double volume = 0., pdmin = 1e10, pdmax = -1e10;
    V2 cen{0, 0}, vel{0, 0};
    for (size_t c = 0; c < m.nc; ++c) {
      double a = o->alpha[i][c];
      volume += a * m.vol[c];
      double pd = o->pd_tc[i][c];
      pdmin = std::min(pdmin, pd);
      pdmax = std::max(pdmax, pd);
      double w = a * m.vol[c];
      cen.x += m.cc[c].x * w;
      cen.y += m.cc[c].y * w;
      vel.x += o->u_tc[0][c] * w;
      vel.y += o->u_tc[1][c] * w;
    }
    st->volume[i] = volume;
    st->mass[i] = volume * o->cfg.density[i];
    st->pd_min[i] = pdmin;
    st->pd_max[i] = pdmax;
    st->cx[i] = cen.x / volume + o->meshpos_stat[0];
    st->cy[i] = cen.y / volume + o->meshpos_stat[1];
    st->volume_in[i] = o->bacc[i][0];
    st->volume_out[i] = o->bacc[i][1];
    st->mass_in[i] = o->bacc[i][2];
    st->mass_out[i] = o->bacc[i][3];
    st->vx[i] = vel.x / volume;
    st->vy[i] = vel.y / volume;
  }
  return HG_OK;
}

// InitHeatSolver (hydro2d.hpp:652-688): uniform temperature_initial; a
// boundary face whose centre lies in the closed box [heat_box_lb,
// heat_box_rt] (vect.hpp:239-246) holds heat_box_temperature, any other a
// zero derivative; the heat solver keeps the dt it is built with.
static void InitHeat(or_ctx* o) {
  const hg_config& cfg = o->cfg;
  const Mesh& m = o->m;
  o->t_tc.assign(m.nc, cfg.temperature_initial);
  o->t_tp = o->t_tc;
  o->hkind.assign(m.nf, 1);
  o->hval.assign(m.nf, 0.);
  for (size_t f = 0; f < m.nf; ++f) {
    if (m.Inner(f)) continue;
    const V2 x = m.fc[f];
    bool inside = true;
    if (x.x < cfg.heat_box_lb[0] || cfg.heat_box_rt[0] < x.x) inside = false;
    if (x.y < cfg.heat_box_lb[1] || cfg.heat_box_rt[1] < x.y) inside = false;
    if (inside) {
      o->hkind[f] = 0;
      o->hval[f] = cfg.heat_box_temperature;
    }
  }
  o->hbdf = DerivCoeffs(0., {-2. * cfg.dt, -cfg.dt, 0.}, cfg.time_second_order_heat ? 0 : 1);
}

// One heat step (hydro2d.hpp:1607-1612): HeatSolver StartStep, one
// MakeIteration (limit 1), FinishStep (heat.hpp:62-80) on
// ConvectionDiffusionScalarImplicit (conv_diff.hpp:117-251) with scaling 1,
// zero source, the fluid's iter_curr flux and the face conductivity
// Interpolate(GetVolumeAveraged(conductivity_i), zero derivative).
static int HeatStep(or_ctx* o) {
  const hg_config& cfg = o->cfg;
  const Mesh& m = o->m;
  const size_t nc = m.nc, nf = m.nf;
  Vec prev(nc);
  for (size_t c = 0; c < nc; ++c) prev[c] = o->t_tc[c] + (o->t_tc[c] - o->t_tp[c]) * 0.;
  // conductivity: cells (hydro2d.hpp:1236-1246, 1428), faces
  Vec kc(nc, 0.), kf(nf);
  for (int i = 0; i < o->np; ++i)
    for (size_t c = 0; c < nc; ++c) kc[c] += cfg.conductivity[i] * o->alpha[i][c];
  Interpolate(m, kc.data(), kf.data(), MakeBc(BC_DERIV, 0.));
  // gradient of the previous iterate with the heat conditions
  Vec tf(nf), gx(nc), gy(nc);
  for (size_t f = 0; f < nf; ++f) {
    if (m.Inner(f)) {
      tf[f] = prev[m.fcm[f]] * (1. - 0.5) + prev[m.fcp[f]] * 0.5;
    } else {
      const int id = m.ValidId(f);
      const size_t c = size_t(id == 0 ? m.fcm[f] : m.fcp[f]);
      const double factor = (id == 0 ? 1. : -1.);
      if (o->hkind[f] == 0) tf[f] = o->hval[f];
      else tf[f] = prev[c] + o->hval[f] * (Norm(m.ToCell(f, id)) * factor);
    }
  }
  Gradient(m, tf.data(), gx.data(), gy.data());
  // face expressions (solver.hpp:201-243, 246-341): convective cam, cap, cak
  // and diffusive dcm, dcp, dck on the two sides of each face
  const Vec& flux = o->fl_ic;
  Vec cam(nf), cap(nf), cak(nf), dcm(nf), dcp(nf), dck(nf);
  for (size_t f = 0; f < nf; ++f) {
    const double F = flux[f], D = kf[f], A_ = m.area[f];
    if (m.Inner(f)) {
      const long cm = m.fcm[f], cp = m.fcp[f];
      double am, ap, ak;
      if (F > 1e-10) {
        am = 1.;
        ap = 0.;
        ak = -Dot(V2{gx[cm], gy[cm]}, m.to0[f]);
      } else if (F < -1e-10) {
        am = 0.;
        ap = 1.;
        ak = -Dot(V2{gx[cp], gy[cp]}, m.to1[f]);
      } else {
        am = 0.5;
        ap = 0.5;
        ak = 0.;
      }
      cam[f] = am * F;
      cap[f] = ap * F;
      cak[f] = ak * F;
      V2 dd{m.to1[f].x - m.to0[f].x, m.to1[f].y - m.to0[f].y};
      const double al = 1. / Norm(dd);
      dcm[f] = ((-al) * (-D)) * A_;
      dcp[f] = (al * (-D)) * A_;
      dck[f] = (0. * (-D)) * A_;
    } else {
      const int id = m.ValidId(f);
      const double factor = (id == 0 ? 1. : -1.);
      const double v = o->hval[f];
      double ccell, dcell;
      if (o->hkind[f] == 0) {
        // value: convective constant v; diffusive -al on the cell, al v
        const double al = 1. / Norm(m.ToCell(f, id)) * factor;
        ccell = 0.;
        cak[f] = v * F;
        dcell = ((-al) * (-D)) * A_;
        dck[f] = ((al * v) * (-D)) * A_;
      } else {
        // derivative: the cell's value + alpha v; diffusive constant v
        const double al = Norm(m.ToCell(f, id)) * factor;
        ccell = 1. * F;
        cak[f] = (al * v) * F;
        dcell = 0.;
        dck[f] = (v * (-D)) * A_;
      }
      cam[f] = id == 0 ? ccell : 0.;
      cap[f] = id == 0 ? 0. : ccell;
      dcm[f] = id == 0 ? dcell : 0.;
      dcp[f] = id == 0 ? 0. : dcell;
    }
  }
  // cell rows (conv_diff.hpp:183-236), merged face sums in W, E, S, N order
  Sys A;
  A.Resize(m.nx, m.ny);
  Vec k(nc);
  const std::vector<double>& co = o->hbdf;
  for (size_t c = 0; c < nc; ++c) {
    double Cc = 0., Ck = 0., Dc = 0., Dk = 0.;
    double Cnb[4] = {0, 0, 0, 0}, Dnb[4] = {0, 0, 0, 0};
    for (int q = 0; q < 4; ++q) {
      const size_t f = m.Face(c, q);
      const double o_ = Mesh::Out(q);
      const bool self_is_m = (q == 1 || q == 3);
      const double cself = self_is_m ? cam[f] : cap[f];
      const double dself = self_is_m ? dcm[f] : dcp[f];
      if (m.Inner(f)) {
        Cc += cself * o_;
        Cnb[q] = (self_is_m ? cap[f] : cam[f]) * o_;
        Dc += dself * o_;
        Dnb[q] = (self_is_m ? dcp[f] : dcm[f]) * o_;
      } else {
        // a boundary face's terms sit on the cell (none for a value's
        // convection or a derivative's diffusion)
        if (o->hkind[f] == 1) Cc += cself * o_;
        if (o->hkind[f] == 0) Dc += dself * o_;
      }
      Ck += cak[f] * o_;
      Dk += dck[f] * o_;
    }
    const double V = m.vol[c];
    const double cen = (Dc / V) + ((Cc / V + co[2]) * 1.);
    double nb[4];
    for (int q = 0; q < 4; ++q) nb[q] = (Dnb[q] / V) + ((Cnb[q] / V) * 1.);
    const double unst = co[0] * o->t_tp[c] + co[1] * o->t_tc[c];
    double kon = ((Dk / V) + ((Ck / V + unst) * 1.)) - 0.;
    const int i = int(c % m.nx), j = int(c / m.nx);
    double ev = kon;
    if (j > 0) ev += prev[c - m.nx] * nb[2];
    if (i > 0) ev += prev[c - 1] * nb[0];
    ev += prev[c] * cen;
    if (i < m.nx - 1) ev += prev[c + 1] * nb[1];
    if (j < m.ny - 1) ev += prev[c + m.nx] * nb[3];
    k[c] = ev;
    A.w[c] = i > 0 ? nb[0] : 0.;
    A.e[c] = i < m.nx - 1 ? nb[1] : 0.;
    A.s[c] = j > 0 ? nb[2] : 0.;
    A.n[c] = j < m.ny - 1 ? nb[3] : 0.;
    A.c[c] = cen / cfg.heat_relaxation_factor;
  }
  Vec corr;
  RunSolver(o, cfg.linear_solver_heat, A, k, corr, false);
  Vec cur(nc);
  for (size_t c = 0; c < nc; ++c) cur[c] = prev[c] + corr[c];
  // FinishStep: time_prev <- time_curr <- iter_curr, "NaN field"
  o->t_tp = o->t_tc;
  o->t_tc = cur;
  for (double v : o->t_tc)
    if (!(v * 0. == 0.)) return HG_ERR_NAN;
  return HG_OK;
}

// hydro::step (hydro2d.hpp:1531-1621), without output.
// FluidSimple::GetAutoTimeStep (fluid.hpp:1191-1206): over cells in order
// and their faces in neighbour order, the smallest |V / flux| of a nonzero
// time_curr flux, from 1e10.
static double AutoTimeStep(const or_ctx* o) {
  const Mesh& m = o->m;
  double dt = 1e10;
  for (size_t c = 0; c < m.nc; ++c)
    for (int k = 0; k < 4; ++k) {
      const double f = o->fl_tc[m.Face(c, k)];
      if (f != 0.) dt = std::min(dt, std::abs(m.vol[c] / f));
    }
  return dt;
}

int or_step(or_ctx* o, hg_step_stats* stats) {
  ScopedTimer t_step(&o->timers, "step");
  hg_step_stats s{};
  std::memset(&s, 0, sizeof(s));
  int rc;
  if (o->cfg.dt_auto) {
    // hydro::step (hydro2d.hpp:1536-1544): fluid dt = dtm * cfl, advection
    // dt = dtm * cfl_advection; the unsteady coefficients follow the new dt
    const double dtm = AutoTimeStep(o);
    o->dt = dtm * o->cfg.cfl;
    o->dt_adv = dtm * o->cfg.cfl_advection;
    o->bdf = DerivCoeffs(0., {-2. * o->dt, -o->dt, 0.}, o->cfg.time_second_order ? 0 : 1);
  }
  {
    ScopedTimer t(&o->timers, "step.fluid");
    rc = or_fluid_start_step(o);
    if (rc) return rc;
    s.pressure_converged = 1;
    if (o->cfg.fluid_enable) {
      while (!or_fluid_is_converged(o)) {
        hg_iter_stats is;
        rc = or_fluid_iterate(o, &is);
        if (rc) return rc;
        s.pressure_solve_iters_total += is.pressure_solve_iters;
        s.convergence_indicator = is.convergence_indicator;
        s.pressure_solve_diff_max = std::max(s.pressure_solve_diff_max, is.pressure_solve_diff);
        s.pressure_converged = s.pressure_converged && is.pressure_converged;
      }
    }
    s.simple_iterations = o->fluid_iter;
    rc = or_fluid_finish_step(o);
    if (rc) return rc;
  }
  if (o->cfg.advection_enable) {
    ScopedTimer t(&o->timers, "step.advection");
    while (o->t_adv < o->t_fluid - 0.5 * o->dt_adv) {
      rc = or_advect_substep(o);
      if (rc) return rc;
      s.advection_substeps += 1;
    }
  }
  if (o->cfg.heat_enable) {
    ScopedTimer t(&o->timers, "step.heat");
    rc = HeatStep(o);
    if (rc) return rc;
  }
  {
    ScopedTimer t(&o->timers, "step.fluid_properties");
    UpdateProperties(o);
    RunningStats(o);
  }
  s.time = o->t_fluid;
  s.dt = o->dt;
  if (stats) *stats = s;
  return HG_OK;
}

double or_timer(const or_ctx* o, const char* region) {
  auto it = o->timers.total.find(region);
  return it == o->timers.total.end() ? -1. : it->second;
}

int or_get_system(or_ctx* o, int which, double* s, double* w, double* c, double* e, double* n,
                  double* k, double* k2) {
  const Sys& a = which == 0 ? o->vel_sys : o->p_sys;
  if (a.size() == 0) return HG_ERR_BAD_ARG;
  size_t N = a.size();
  std::copy(a.s.begin(), a.s.end(), s);
  std::copy(a.w.begin(), a.w.end(), w);
  std::copy(a.c.begin(), a.c.end(), c);
  std::copy(a.e.begin(), a.e.end(), e);
  std::copy(a.n.begin(), a.n.end(), n);
  if (which == 0) {
    std::copy(o->vel_k[0].begin(), o->vel_k[0].end(), k);
    if (k2) std::copy(o->vel_k[1].begin(), o->vel_k[1].end(), k2);
  } else {
    std::copy(a.k.begin(), a.k.begin() + N, k);
  }
  return HG_OK;
}

int or_stencil_solve(int nx, int ny, const double* s, const double* w, const double* c,
                     const double* e, const double* n, const double* k, double* x,
                     const hg_solve_opts* opts, hg_solve_stats* stats) {
  if (nx < 1 || ny < 1) return HG_ERR_BAD_ARG;
  Sys a;
  a.Resize(nx, ny);
  size_t N = a.size();
  std::copy(s, s + N, a.s.begin());
  std::copy(w, w + N, a.w.begin());
  std::copy(c, c + N, a.c.begin());
  std::copy(e, e + N, a.e.begin());
  std::copy(n, n + N, a.n.begin());
  std::copy(k, k + N, a.k.begin());
  Vec xv;
  SolveResult r;
  switch (opts->solver) {
    case HG_SOLVER_LU: SolveLu(a, a.k, xv); break;
    case HG_SOLVER_LU_RB: SolveLuRb(a, a.k, xv); break;
    case HG_SOLVER_GAUSS_SEIDEL:
      r = SolveGs(a, a.k, xv, opts->tolerance, opts->max_iters, opts->relaxation_factor);
      break;
    case HG_SOLVER_JACOBI:
      r = SolveJacobi(a, a.k, xv, opts->tolerance, opts->max_iters, opts->relaxation_factor);
      break;
    case HG_SOLVER_LU_RELAXED:
      r = SolveLuRelaxed(a, a.k, xv, opts->tolerance, opts->max_iters, opts->relaxation_factor);
      break;
    case HG_SOLVER_CG:
      r = SolveCg(a, a.k, xv, opts->tolerance, opts->max_iters, false);
      break;
    default:
      return HG_ERR_UNSUPPORTED;
  }
  std::copy(xv.begin(), xv.end(), x);
  if (stats) {
    stats->iterations = r.iters;
    stats->final_diff = r.diff;
    stats->converged = r.converged ? 1 : 0;
    double bn = 0., rn = 0.;
    for (size_t i = 0; i < N; ++i) {
      double res = a.Eval(i, xv.data());
      rn += res * res;
      bn += a.k[i] * a.k[i];
    }
    stats->rel_residual = bn > 0 ? std::sqrt(rn / bn) : 0.;
    stats->ms = 0.;
  }
  return HG_OK;
}

int or_poisson_solve(int nx, int ny, const double* kx, const double* ky, const double* rhs,
                     double* p, const hg_solve_opts* opts, hg_solve_stats* stats) {
  Sys a;
  a.Resize(nx, ny);
  size_t N = a.size();
  size_t nfx = size_t(nx + 1) * ny;
  (void)nfx;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      size_t r = size_t(i) + size_t(nx) * j;
      double kw = kx[size_t(i) + size_t(nx + 1) * j], ke = kx[size_t(i + 1) + size_t(nx + 1) * j];
      double ks = ky[size_t(i) + size_t(nx) * j], kn = ky[size_t(i) + size_t(nx) * (j + 1)];
      a.w[r] = -kw;
      a.e[r] = -ke;
      a.s[r] = -ks;
      a.n[r] = -kn;
      a.c[r] = ((kw + ke) + ks) + kn;
      a.k[r] = -rhs[r];
    }
  Vec xv;
  SolveResult r = SolveCg(a, a.k, xv, opts->tolerance, opts->max_iters, true);
  std::copy(xv.begin(), xv.end(), p);
  if (stats) {
    stats->iterations = r.iters;
    stats->rel_residual = r.diff;
    stats->final_diff = r.diff;
    stats->converged = r.converged ? 1 : 0;
    stats->ms = 0.;
  }
  (void)N;
  return HG_OK;
}

// ---- CPU baselines of the pressure solve to a relative residual --------------
// The face-form system A p = b of or_poisson_solve (b projected to zero mean),
// solved from p = 0 until ||b - A p||_2 <= tol ||b||_2:
//  method 0: the reference's pressure algorithm, lexicographic SOR
//            (GaussSeidel::Solve, linear.hpp:692-709: value = -(k + sum) / c,
//            x += omega (value - x)), single-threaded, the residual checked
//            every `check_every` sweeps instead of the max-update test;
//  method 1: Jacobi-preconditioned CG with every loop and dot product under
//            OpenMP (the strong CPU baseline; no reference counterpart).
// Stops after max_iters sweeps / iterations.  *seconds: wall time of the
// solve (residual checks included).
int or_poisson_baseline(int nx, int ny, const double* kx, const double* ky, const double* rhs,
                        int method, double omega, double tol, int max_iters, int check_every,
                        double* p, double* seconds, int* iters, double* rel_res) {
  const long N = long(nx) * ny;
  Vec w(N), e(N), s_(N), n(N), c(N), b(N);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const long r = i + long(nx) * j;
      w[r] = -kx[size_t(i) + size_t(nx + 1) * j];
      e[r] = -kx[size_t(i + 1) + size_t(nx + 1) * j];
      s_[r] = -ky[size_t(i) + size_t(nx) * j];
      n[r] = -ky[size_t(i) + size_t(nx) * (j + 1)];
      c[r] = -(((w[r] + e[r]) + s_[r]) + n[r]);
      b[r] = rhs[r];
    }
  double mean = 0.;
  for (double v : b) mean += v;
  mean /= double(N);
  for (double& v : b) v -= mean;
  auto apply = [&](const Vec& x, Vec& y) {
#pragma omp parallel for
    for (long r = 0; r < N; ++r) {
      const int i = int(r % nx), j = int(r / nx);
      double v = c[r] * x[r];
      if (j > 0) v += s_[r] * x[r - nx];
      if (i > 0) v += w[r] * x[r - 1];
      if (i < nx - 1) v += e[r] * x[r + 1];
      if (j < ny - 1) v += n[r] * x[r + nx];
      y[r] = v;
    }
  };
  auto dot = [&](const Vec& a, const Vec& bb) {
    double t = 0.;
#pragma omp parallel for reduction(+ : t)
    for (long r = 0; r < N; ++r) t += a[r] * bb[r];
    return t;
  };
  const double bnorm = std::sqrt(dot(b, b));
  Vec x(N, 0.), q(N);
  auto relres = [&]() {
    apply(x, q);
    double t = 0.;
#pragma omp parallel for reduction(+ : t)
    for (long r = 0; r < N; ++r) t += (b[r] - q[r]) * (b[r] - q[r]);
    return bnorm > 0 ? std::sqrt(t) / bnorm : 0.;
  };
  const auto t0 = std::chrono::steady_clock::now();
  int it = 0;
  double rel = bnorm > 0 ? 1. : 0.;
  if (method == 0) {
    while (rel > tol && it < max_iters) {
      for (long r = 0; r < N; ++r) {
        const int i = int(r % nx), j = int(r / nx);
        double sum = 0.;
        if (j > 0) sum += s_[r] * x[r - nx];
        if (i > 0) sum += w[r] * x[r - 1];
        if (i < nx - 1) sum += e[r] * x[r + 1];
        if (j < ny - 1) sum += n[r] * x[r + nx];
        const double value = -(-b[r] + sum) / c[r];
        x[r] += (value - x[r]) * omega;
      }
      ++it;
      if (it % check_every == 0 || it == max_iters) rel = relres();
    }
  } else {
    Vec r(b), z(N), d(N);
#pragma omp parallel for
    for (long i = 0; i < N; ++i) z[i] = r[i] / c[i];
    d = z;
    double rz = dot(r, z);
    while (rel > tol && it < max_iters) {
      apply(d, q);
      const double alpha = rz / dot(d, q);
#pragma omp parallel for
      for (long i = 0; i < N; ++i) {
        x[i] += alpha * d[i];
        r[i] -= alpha * q[i];
      }
      ++it;
      rel = std::sqrt(dot(r, r)) / bnorm;
      if (rel <= tol) break;
#pragma omp parallel for
      for (long i = 0; i < N; ++i) z[i] = r[i] / c[i];
      const double rz2 = dot(r, z);
      const double beta = rz2 / rz;
      rz = rz2;
#pragma omp parallel for
      for (long i = 0; i < N; ++i) d[i] = z[i] + beta * d[i];
    }
    rel = relres();
  }
  *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  *iters = it;
  *rel_res = rel;
  std::copy(x.begin(), x.end(), p);
  return 0;
}

int or_geometry(const or_ctx* o, double* cx, double* cy, double* cv, double* fx, double* fy,
                double* fa, double* sx, double* sy) {
  const Mesh& m = o->m;
  for (size_t c = 0; c < m.nc; ++c) {
    cx[c] = m.cc[c].x;
    cy[c] = m.cc[c].y;
    cv[c] = m.vol[c];
  }
  for (size_t f = 0; f < m.nf; ++f) {
    fx[f] = m.fc[f].x;
    fy[f] = m.fc[f].y;
    fa[f] = m.area[f];
    sx[f] = m.surf[f].x;
    sy[f] = m.surf[f].y;
  }
  return HG_OK;
}

void or_bdf_coeffs(double dt, int second_order, double out[3]) {
  std::vector<double> c = DerivCoeffs(0., {-2. * dt, -dt, 0.}, second_order ? 0 : 1);
  out[0] = c[0];
  out[1] = c[1];
  out[2] = c[2];
}

}  // extern "C"
