// sanitize_driver.cpp -- the host-only C++ of this repository under
// AddressSanitizer and UndefinedBehaviorSanitizer (built and run by
// tests/test_sanitizers.py with g++ -fsanitize=address,undefined).
//
// Exercises:
//  * the CPU oracle (oracle/hydro_oracle.cpp): a few steps of every case
//    family the tests use (cavity, two-phase dam, inlet/outlet channel,
//    SIMPLER, heat, dt_auto, every row solver and CG), the stencil and
//    Poisson solves and the CPU solve baselines;
//  * the 3D oracle (oracle/hydro_oracle3d.cpp): steps of a cavity with each
//    of its solvers and of a two-phase dam;
//  * the configuration reader (hydro_amd/csrc/config.cpp, 2D and 3D) on
//    valid and malformed scripts;
//  * the output keys and the .vts writer (hydro_amd/csrc/output.cpp).
// output.cpp's session entry points read a context through the C-ABI; the
// sessions are not run here, so those calls resolve to the refusals below.

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../include/hydro_gpu.h"
#include "../include/hydro_output.h"
#include "../oracle/hydro_oracle.h"
#include "../oracle/hydro_oracle3d.h"

extern "C" {
// Context entry points output.cpp links against (libhydrogpu.so's
// context.hip provides them); no context exists in this driver.
hg_status hg_get_field(hg_ctx*, int32_t, double*, size_t) { return HG_ERR_UNSUPPORTED; }
hg_status hg_get_config(const hg_ctx*, hg_config*) { return HG_ERR_UNSUPPORTED; }
hg_status hg_get_progress(hg_ctx*, hg_progress*) { return HG_ERR_UNSUPPORTED; }
hg_status hg_get_iteration_history(hg_ctx*, double*, int32_t, int32_t*) {
  return HG_ERR_UNSUPPORTED;
}
hg_status hg_calc_stats(hg_ctx*, hg_stats*) { return HG_ERR_UNSUPPORTED; }
const char* hg_last_error(const hg_ctx*) { return "no context in this driver"; }
}

namespace {

int failures = 0;

void Check(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    ++failures;
  }
}

hg_config Parse(const std::string& text) {
  hg_config cfg;
  hg_config_default(&cfg);
  char err[256] = {0};
  const hg_status rc = hg_config_parse(&cfg, text.c_str(), err, sizeof err);
  Check(rc == HG_OK, ("parse: " + std::string(err)).c_str());
  return cfg;
}

void RunOracle(const char* name, const std::string& text, int steps) {
  hg_config cfg = Parse(text);
  or_ctx* o = nullptr;
  char err[256] = {0};
  if (or_create(&cfg, &o, err, sizeof err) != 0) {
    std::fprintf(stderr, "FAILED: or_create %s: %s\n", name, err);
    ++failures;
    return;
  }
  for (int s = 0; s < steps; ++s) {
    hg_step_stats st{};
    Check(or_step(o, &st) == 0, name);
  }
  hg_stats stats{};
  Check(or_calc_stats(o, &stats) == 0, name);
  const int nc = or_num_cells(o);
  std::vector<double> p(static_cast<size_t>(nc));
  Check(or_get_field(o, HG_FIELD_PRESSURE, p.data(), p.size()) == 0, name);
  for (double v : p) Check(std::isfinite(v), name);
  or_destroy(o);
  std::printf("oracle %s: %d steps\n", name, steps);
}

const char* kCavity =
    "set int Nx 16\nset int Ny 16\nset double dt 1e-3\nset double viscosity_0 0.01\n"
    "set string condition_top \"wall 1 0\"\n";
const char* kDam =
    "set vect A (0, 0)\nset vect B (2, 1)\nset int Nx 16\nset int Ny 8\nset int num_phases 2\n"
    "set double density_0 1\nset double density_1 1000\nset double viscosity_0 1e-5\n"
    "set double viscosity_1 1e-3\nset vect gravity (0, -9.81)\nset vect A1 (0, 0)\n"
    "set vect B1 (0.5, 0.5)\nset double dt 1e-4\n";
const char* kChannel =
    "set vect A (0, 0)\nset vect B (2, 1)\nset int Nx 16\nset int Ny 8\nset int num_phases 2\n"
    "set double density_1 5\nset double viscosity_0 0.01\nset double viscosity_1 0.02\n"
    "set vect gravity (0, -1)\nset vect A1 (0, 0)\nset vect B1 (0.4, 0.3)\nset double dt 2e-3\n"
    "set string condition_left \"inlet 1 0\"\nset string condition_right \"outlet\"\n";

}  // namespace

int main() {
  const std::string cav(kCavity), dam(kDam), chan(kChannel);
  RunOracle("cavity gauss_seidel+lu", cav, 3);
  RunOracle("cavity jacobi", cav + "set string linear_solver_pressure jacobi\n"
                                   "set double lu_relaxed_relaxation_factor 0.9\n", 2);
  RunOracle("cavity lu_relaxed both", cav + "set string linear_solver_pressure lu_relaxed\n"
                                            "set string linear_solver_velocity lu_relaxed\n", 2);
  RunOracle("cavity cg+lu_rb", cav + "set string linear_solver_pressure cg\n"
                                     "set string linear_solver_velocity lu_rb\n", 3);
  RunOracle("cavity simpler", cav + "set bool simpler 1\n", 2);
  RunOracle("dam", dam, 3);
  RunOracle("dam heat", dam + "set bool heat_enable 1\nset vect heat_box_lb (0, 0)\n"
                              "set vect heat_box_rt (0, 1)\nset double heat_box_temperature 2\n",
            2);
  RunOracle("dam dt_auto", dam + "set bool dt_auto 1\nset double cfl 0.02\n"
                                 "set double cfl_advection 0.01\n"
                                 "set vect initial_velocity (0.2, 0.1)\n", 2);
  RunOracle("channel", chan, 3);

  // the 3D oracle
  {
    const std::string w3 = "set vect initial_velocity (0, 0, 0)\nset int Nx 6\nset int Ny 5\n"
                           "set int Nz 4\nset vect A (0, 0, 0)\nset vect B (1, 0.8, 0.6)\n"
                           "set double dt 1e-3\nset double lu_relaxed_relaxation_factor 0.9\n";
    const char* cases[] = {
        "set string condition_top \"wall 1 0 0.5\"\nset string linear_solver_pressure jacobi\n"
        "set string linear_solver_velocity jacobi\n",
        "set string condition_far \"wall 0 1 0\"\n",  // gauss_seidel + lu
        "set string linear_solver_pressure cg\nset string linear_solver_velocity jacobi\n"
        "set double convergence_tolerance 1e-5\n",
        "set int num_phases 2\nset double density_1 1000\nset vect gravity (0, 0, -9.81)\n"
        "set vect A1 (0, 0, 0)\nset vect B1 (0.5, 0.8, 0.3)\nset double dt 1e-4\n"};
    for (const char* extra : cases) {
      hg3_config c3;
      hg3_config_default(&c3);
      char err[256] = {0};
      Check(hg3_config_parse(&c3, (w3 + extra).c_str(), err, sizeof err) == HG_OK, "parse3");
      or3_ctx* o = nullptr;
      Check(or3_create(&c3, &o, err, sizeof err) == 0, "or3_create");
      if (!o) continue;
      for (int s = 0; s < 2; ++s) {
        hg_step_stats st{};
        Check(or3_step(o, &st) == 0, "or3_step");
      }
      std::vector<double> u(size_t(or3_num_cells(o)));
      Check(or3_get_field(o, HG_FIELD_VELOCITY_Z, u.data(), u.size()) == 0, "or3_get_field");
      for (double v : u) Check(std::isfinite(v), "3D field finite");
      or3_destroy(o);
    }
    hg3_config c3;
    hg3_config_default(&c3);
    char err[64];
    (void)hg3_config_parse(&c3, "set int Nz\nset vect B (1,\nset string condition_far \"wall\"\n",
                           err, sizeof err);
    std::printf("oracle 3D: %d cases\n", int(sizeof cases / sizeof cases[0]));
  }

  // stencil and Poisson solves, and the CPU baselines
  {
    const int nx = 12, ny = 10, n = nx * ny;
    std::vector<double> s(n), w(n), c(n), e(n), nn(n), k(n), x(n);
    for (int r = 0; r < n; ++r) {
      const int i = r % nx, j = r / nx;
      s[r] = j > 0 ? -1. : 0.;
      w[r] = i > 0 ? -1. : 0.;
      e[r] = i < nx - 1 ? -1. : 0.;
      nn[r] = j < ny - 1 ? -1. : 0.;
      c[r] = 4.5;
      k[r] = std::sin(0.3 * r);
    }
    for (int solver : {HG_SOLVER_LU, HG_SOLVER_LU_RB, HG_SOLVER_GAUSS_SEIDEL, HG_SOLVER_JACOBI,
                       HG_SOLVER_LU_RELAXED, HG_SOLVER_CG}) {
      hg_solve_opts o{solver, 1e-8, 200, solver == HG_SOLVER_JACOBI ? 0.8 : 1.2, 0};
      hg_solve_stats st{};
      Check(or_stencil_solve(nx, ny, s.data(), w.data(), c.data(), e.data(), nn.data(), k.data(),
                             x.data(), &o, &st) == 0,
            "or_stencil_solve");
    }
    std::vector<double> kx(size_t(nx + 1) * ny, 1.), ky(size_t(nx) * (ny + 1), 1.), rhs(n);
    for (int j = 0; j < ny; ++j) {
      kx[size_t(nx + 1) * j] = kx[size_t(nx + 1) * j + nx] = 0.;
    }
    for (int i = 0; i < nx; ++i) ky[i] = ky[size_t(nx) * ny + i] = 0.;
    double mean = 0.;
    for (int r = 0; r < n; ++r) mean += (rhs[r] = std::cos(0.7 * r));
    for (int r = 0; r < n; ++r) rhs[r] -= mean / n;
    hg_solve_opts o{HG_SOLVER_CG, 1e-10, 1000, 1., 0};
    hg_solve_stats st{};
    Check(or_poisson_solve(nx, ny, kx.data(), ky.data(), rhs.data(), x.data(), &o, &st) == 0,
          "or_poisson_solve");
    for (int method : {0, 1}) {
      double sec = 0., rel = 0.;
      int it = 0;
      Check(or_poisson_baseline(nx, ny, kx.data(), ky.data(), rhs.data(), method, 1.9, 1e-8,
                                20000, 10, x.data(), &sec, &it, &rel) == 0 &&
                rel <= 1e-8,
            "or_poisson_baseline");
    }
  }

  // configuration reader: valid scripts, then malformed lines that must be
  // refused without reading outside their text
  {
    hg_config cfg = Parse(cav + "set string precision_gpu mixed\nset vect meshvel (0.1, 0)\n"
                                "set string linear_solver_velocity lu_rb\n"
                                "set string condition_bottom \"wall 0.5 -1\"\n");
    Check(cfg.precision_gpu == HG_PREC_MIXED, "precision_gpu");
    const char* bad[] = {"set int Nx", "set double dt x", "set vect A (1,", "set string",
                         "set vect B (1, 2, 3, 4, 5, 6, 7)", "set int Nx 99999999999999999999",
                         "set string linear_solver_pressure \"unterminated",
                         "set string condition_top \"wall 1\"", "set int", "set"};
    for (const char* line : bad) {
      hg_config c2;
      hg_config_default(&c2);
      char err[64];
      (void)hg_config_parse(&c2, line, err, sizeof err);  // refusal or no effect
    }
    for (const char* name : {"lu", "lu_relaxed", "gauss_seidel", "jacobi", "lu_rb", "cg",
                             "fcg_mg", "fcg_learned", "", "nonsense"})
      (void)hg_solver_from_name(name);
  }

  // output keys and the .vts writer
  {
    hg_output_config oc;
    hg_output_config_default(&oc);
    char err[128] = {0};
    Check(hg_output_config_parse(&oc, "set string name run\nset int output_factor_x 2\n"
                                      "set bool output_divergence 1\nset double T 0.5\n",
                                 err, sizeof err) == HG_OK,
          "output parse");
    const int nx = 5, ny = 3;
    std::vector<double> xn(nx + 1), yn(ny + 1), a(size_t(nx) * ny), b(size_t(nx) * ny);
    for (int i = 0; i <= nx; ++i) xn[i] = 0.2 * i;
    for (int j = 0; j <= ny; ++j) yn[j] = 0.25 * j;
    for (size_t q = 0; q < a.size(); ++q) {
      a[q] = 1e-7 * q - 3.;
      b[q] = std::exp(-double(q));
    }
    const char* names[] = {"pressure", "volume_fraction_0"};
    const double* vals[] = {a.data(), b.data()};
    const std::string path = "sanitize_frame.vts";
    Check(hg_vts_write(path.c_str(), nx, ny, xn.data(), yn.data(), 0, nullptr, nullptr, 2, names,
                       vals) == HG_OK,
          "hg_vts_write");
    std::remove(path.c_str());
  }

  std::printf(failures ? "SANITIZE_FAILED %d\n" : "SANITIZE_OK\n", failures);
  return failures ? 1 : 0;
}
