// config.cpp -- the .hydroconf interface of the time step (host code).
//
// hg_config_default() holds the values of the reference's
// examples/general.hydroconf; hg_config_parse() applies
// `set <type> <name> <value>` and `del <name>` lines the way the reference
// console does (source/control/console.cpp:276-342: typed maps, later lines
// override earlier ones, unknown names are stored and ignored).

#include <cctype>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/hydro_gpu.h"
#include "../../include/hydro_gpu3d.h"

namespace {

bool ReadVect(std::istream& in, std::vector<double>& out) {
  // Reference syntax "(a, b, c)" (common/linear_algebra.hpp:775-799).
  out.clear();
  char p = 0;
  in >> p;
  if (p != '(') return false;
  in >> std::ws;
  if (in.peek() == ')') {
    in.get();
    return true;
  }
  while (p != ')') {
    double a;
    in >> a;
    if (in.fail()) return false;
    out.push_back(a);
    in >> p;
    if (in.fail()) return false;
  }
  return true;
}

std::string RestOfLine(std::istream& in) {
  std::string s;
  std::getline(in, s);
  size_t b = s.find_first_not_of(" \t");
  if (b == std::string::npos) return "";
  s = s.substr(b);
  size_t e = s.find_last_not_of(" \t\r");
  s = s.substr(0, e + 1);
  if (s.size() >= 2 && s.front() == '"' && s.back() == '"') s = s.substr(1, s.size() - 2);
  return s;
}

bool Set2(double dst[2], const std::vector<double>& v) {
  if (v.size() < 2) return false;
  dst[0] = v[0];
  dst[1] = v[1];
  return true;
}

}  // namespace

extern "C" {

void hg_config_default(hg_config* c) {
  std::memset(c, 0, sizeof(*c));
  c->abi_version = HG_ABI_VERSION;
  c->nx = 100;
  c->ny = 100;
  c->a[0] = 0; c->a[1] = 0;
  c->b[0] = 1; c->b[1] = 1;
  c->dt = 0.01;
  c->num_phases = 1;
  for (int i = 0; i < HG_MAX_PHASES; ++i) {
    c->density[i] = 1;
    c->viscosity[i] = 1;
  }
  c->sigma = 0.;
  for (int s = 0; s < 4; ++s) c->bc_kind[s] = HG_BC_WALL;
  c->b1[1] = 0; c->b2[1] = 0;  // A1=B1=(0,0), A2=B2=(0,0)
  c->initial_volume_fraction_smooth_times = 2;
  c->density_smooth_times = 2;
  c->viscosity_smooth_times = 2;
  c->force_smooth_times = 0;
  c->fluid_enable = 1;
  c->advection_enable = 1;
  c->dt_auto = 0;
  c->cfl = 0.5;
  c->cfl_advection = 0.5;
  c->velocity_relaxation_factor = 0.8;
  c->pressure_relaxation_factor = 0.9;
  c->rhie_chow_factor = 1;
  c->convergence_tolerance = 1e-2;
  c->num_iterations_limit = 10;
  c->time_second_order = 1;
  c->simpler = 0;
  c->force_geometric_average = 0;
  c->guess_extrapolation = 0;
  c->linear_solver_velocity = HG_SOLVER_LU;
  c->linear_solver_pressure = HG_SOLVER_GAUSS_SEIDEL;
  c->lu_relaxed_tolerance = 1e-3;
  c->lu_relaxed_num_iters_limit = 1000;
  c->lu_relaxed_relaxation_factor = 1.9;
  c->advection_dt_factor = 0.1;
  c->tvd_split = 0;
  c->sharp = 0.;
  c->heat_enable = 0;
  c->linear_solver_heat = HG_SOLVER_LU;
  for (int i = 0; i < HG_MAX_PHASES; ++i) c->conductivity[i] = 1;
  c->temperature_initial = 0;
  c->heat_box_temperature = 0;
  c->heat_relaxation_factor = 1.;
  c->time_second_order_heat = 1;
  c->pressure_tolerance_gpu = 1e-8;
  c->pressure_max_iters_gpu = 20000;
  c->precision_gpu = HG_PREC_FP64;
  c->learned_levels_gpu = 4;
  c->learned_channels_gpu = 16;
  c->pressure_warm_start_gpu = 1;
  c->meshvel_output = 1;
}

int32_t hg_solver_from_name(const char* name) {
  std::string n(name ? name : "");
  if (n == "lu") return HG_SOLVER_LU;
  if (n == "lu_relaxed") return HG_SOLVER_LU_RELAXED;
  if (n == "gauss_seidel") return HG_SOLVER_GAUSS_SEIDEL;
  if (n == "jacobi") return HG_SOLVER_JACOBI;
  if (n == "lu_rb") return HG_SOLVER_LU_RB;
  if (n == "cg") return HG_SOLVER_CG;
  if (n == "fcg_learned") return HG_SOLVER_FCG_LEARNED;
  if (n == "fcg_mg") return HG_SOLVER_FCG_MG;
  return -1;
}

hg_status hg_config_parse(hg_config* c, const char* text, char* err, size_t errlen) {
  auto fail = [&](int line_no, const std::string& msg) {
    if (err && errlen) std::snprintf(err, errlen, "line %d: %s", line_no, msg.c_str());
    return HG_ERR_PARSE;
  };
  std::istringstream all(text ? text : "");
  std::string line;
  int line_no = 0;
  while (std::getline(all, line)) {
    ++line_no;
    size_t hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    std::istringstream in(line);
    std::string cmd;
    if (!(in >> cmd)) continue;
    if (cmd == "del") {
      std::string name;
      in >> name;
      if (name == "pressure_fixed_point") c->has_pressure_fixed_point = 0;
      if (name == "initial_sin_n") c->has_initial_sin = 0;
      if (name == "antidiffusion_factor") c->unsupported_flags &= ~HG_UNSUP_ANTIDIFFUSION;
      continue;
    }
    if (cmd != "set") continue;  // ae, run, init, start, ... are console verbs
    std::string type, name;
    if (!(in >> type >> name)) return fail(line_no, "incomplete set");
    if (type == "int") {
      long v;
      if (!(in >> v)) return fail(line_no, "bad int for " + name);
      int32_t iv = int32_t(v);
      if (name == "Nx") c->nx = iv;
      else if (name == "Ny") c->ny = iv;
      else if (name == "num_phases") c->num_phases = iv;
      else if (name == "initial_volume_fraction_smooth_times") c->initial_volume_fraction_smooth_times = iv;
      else if (name == "density_smooth_times") c->density_smooth_times = iv;
      else if (name == "viscosity_smooth_times") c->viscosity_smooth_times = iv;
      else if (name == "force_smooth_times") c->force_smooth_times = iv;
      else if (name == "num_iterations_limit") c->num_iterations_limit = iv;
      else if (name == "lu_relaxed_num_iters_limit") c->lu_relaxed_num_iters_limit = iv;
      else if (name == "pressure_max_iters_gpu") c->pressure_max_iters_gpu = iv;
      else if (name == "learned_levels_gpu") c->learned_levels_gpu = iv;
      else if (name == "learned_channels_gpu") c->learned_channels_gpu = iv;
      else if (name == "pressure_warm_start_gpu") c->pressure_warm_start_gpu = iv;
      else if (name == "fcg_negate_precond_gpu") c->fcg_negate_precond_gpu = iv;
    } else if (type == "double") {
      double v;
      if (!(in >> v)) return fail(line_no, "bad double for " + name);
      if (name == "dt") c->dt = v;
      else if (name.rfind("density_", 0) == 0 && name.size() == 9 && isdigit(name[8]) && name[8] - '0' < HG_MAX_PHASES)
        c->density[name[8] - '0'] = v;
      else if (name.rfind("viscosity_", 0) == 0 && name.size() == 11 && isdigit(name[10]) && name[10] - '0' < HG_MAX_PHASES)
        c->viscosity[name[10] - '0'] = v;
      else if (name.rfind("initial_volume_fraction_", 0) == 0 && name.size() == 25 && isdigit(name[24]) && name[24] - '0' < HG_MAX_PHASES)
        c->initial_volume_fraction[name[24] - '0'] = v;
      else if (name == "sigma") c->sigma = v;
      else if (name == "IR") c->ir = v;
      else if (name == "IR2") c->ir2 = v;
      else if (name == "velocity_relaxation_factor") c->velocity_relaxation_factor = v;
      else if (name == "pressure_relaxation_factor") c->pressure_relaxation_factor = v;
      else if (name == "rhie_chow_factor") c->rhie_chow_factor = v;
      else if (name == "convergence_tolerance") c->convergence_tolerance = v;
      else if (name == "guess_extrapolation") c->guess_extrapolation = v;
      else if (name == "pressure_fixed_value") c->pressure_fixed_value = v;
      else if (name == "lu_relaxed_tolerance") c->lu_relaxed_tolerance = v;
      else if (name == "lu_relaxed_relaxation_factor") c->lu_relaxed_relaxation_factor = v;
      else if (name == "advection_dt_factor") c->advection_dt_factor = v;
      else if (name == "sharp") c->sharp = v;
      else if (name == "cfl") c->cfl = v;
      else if (name.rfind("conductivity_", 0) == 0 && name.size() == 14 && isdigit(name[13]) && name[13] - '0' < HG_MAX_PHASES)
        c->conductivity[name[13] - '0'] = v;
      else if (name == "temperature_initial") c->temperature_initial = v;
      else if (name == "heat_box_temperature") c->heat_box_temperature = v;
      else if (name == "heat_relaxation_factor") c->heat_relaxation_factor = v;
      else if (name == "cfl_advection") c->cfl_advection = v;
      else if (name == "initial_sin_lambda") c->initial_sin_lambda = v;
      else if (name == "initial_sin_phase") c->initial_sin_phase = v;
      else if (name == "pressure_tolerance_gpu") c->pressure_tolerance_gpu = v;
      else if (name == "chem_intensity") {
        if (v != 0.) c->unsupported_flags |= HG_UNSUP_CHEMISTRY;
        else c->unsupported_flags &= ~HG_UNSUP_CHEMISTRY;
      } else if (name == "antidiffusion_factor") c->unsupported_flags |= HG_UNSUP_ANTIDIFFUSION;
    } else if (type == "bool") {
      int v;
      if (!(in >> v)) return fail(line_no, "bad bool for " + name);
      auto flag = [&](int bit) {
        if (v) c->unsupported_flags |= bit;
        else c->unsupported_flags &= ~bit;
      };
      if (name == "fluid_enable") c->fluid_enable = v;
      else if (name == "advection_enable") c->advection_enable = v;
      else if (name == "time_second_order") c->time_second_order = v;
      else if (name == "simpler") c->simpler = v;
      else if (name == "fcg_negate_precond_gpu") c->fcg_negate_precond_gpu = v;
      else if (name == "force_geometric_average") c->force_geometric_average = v;
      else if (name == "tvd_split") c->tvd_split = v;
      else if (name == "meshvel_output") c->meshvel_output = v;
      else if (name == "dt_auto") c->dt_auto = v;
      else if (name == "compressible_enable") flag(HG_UNSUP_COMPRESSIBLE);
      else if (name == "heat_enable") c->heat_enable = v;
      else if (name == "time_second_order_heat") c->time_second_order_heat = v;
      else if (name == "radiation_enable") flag(HG_UNSUP_RADIATION);
      else if (name.rfind("enable_settling_", 0) == 0) flag(HG_UNSUP_SETTLING);
      else if (name == "initial_pois" || name == "deforming_velocity") flag(HG_UNSUP_OTHER);
    } else if (type == "string") {
      std::string v = RestOfLine(in);
      if (name.rfind("condition_", 0) == 0) {
        int side = -1;
        if (name == "condition_left") side = HG_SIDE_LEFT;
        else if (name == "condition_right") side = HG_SIDE_RIGHT;
        else if (name == "condition_bottom") side = HG_SIDE_BOTTOM;
        else if (name == "condition_top") side = HG_SIDE_TOP;
        if (side < 0) continue;  // condition_close / condition_far (3D)
        std::istringstream cs(v);
        std::string kind;
        cs >> kind;
        if (kind == "wall") {
          double vx = 0, vy = 0;
          cs >> vx >> vy;
          c->bc_kind[side] = HG_BC_WALL;
          c->bc_velocity[side][0] = vx;
          c->bc_velocity[side][1] = vy;
        } else if (kind == "inlet") {
          double vx = 0, vy = 0;
          cs >> vx >> vy;
          c->bc_kind[side] = HG_BC_INLET;
          c->bc_velocity[side][0] = vx;
          c->bc_velocity[side][1] = vy;
        } else if (kind == "outlet") {
          c->bc_kind[side] = HG_BC_OUTLET;
          c->bc_velocity[side][0] = 0.;
          c->bc_velocity[side][1] = 0.;
        } else {
          return fail(line_no, "unknown boundary condition '" + kind + "'");
        }
      } else if (name == "linear_solver_velocity" || name == "linear_solver_pressure" ||
                 name == "linear_solver_heat") {
        int32_t s = hg_solver_from_name(v.c_str());
        if (s < 0) return fail(line_no, "Unknown linear solver '" + v + "'");
        if (name == "linear_solver_velocity") c->linear_solver_velocity = s;
        else if (name == "linear_solver_heat") c->linear_solver_heat = s;
        else c->linear_solver_pressure = s;
      } else if (name == "advection_solver") {
        if (v == "tvd") c->unsupported_flags &= ~HG_UNSUP_ADVECTION;
        else c->unsupported_flags |= HG_UNSUP_ADVECTION;
      } else if (name == "precision_gpu") {
        if (v == "fp64") c->precision_gpu = HG_PREC_FP64;
        else if (v == "mixed") c->precision_gpu = HG_PREC_MIXED;
        else return fail(line_no, "precision_gpu must be fp64 or mixed");
      } else if (name == "img_init" || name == "imgu_init" || name == "imgv_init") {
        c->unsupported_flags |= HG_UNSUP_IMAGE_INIT;
      } else if (name == "meshvel_auto") {
        c->unsupported_flags |= HG_UNSUP_OTHER;
      }
    } else if (type == "vect") {
      std::vector<double> v;
      if (!ReadVect(in, v)) return fail(line_no, "bad vect for " + name);
      bool ok = true;
      if (name == "A") ok = Set2(c->a, v);
      else if (name == "B") ok = Set2(c->b, v);
      else if (name == "A1") ok = Set2(c->a1, v);
      else if (name == "B1") ok = Set2(c->b1, v);
      else if (name == "A2") ok = Set2(c->a2, v);
      else if (name == "B2") ok = Set2(c->b2, v);
      else if (name == "IC") ok = Set2(c->ic, v);
      else if (name == "IC2") ok = Set2(c->ic2, v);
      else if (name == "gravity") ok = Set2(c->gravity, v);
      else if (name == "force") ok = Set2(c->force, v);
      else if (name == "initial_velocity") ok = Set2(c->initial_velocity, v);
      else if (name == "meshvel") ok = Set2(c->meshvel, v);
      else if (name == "heat_box_lb") ok = Set2(c->heat_box_lb, v);
      else if (name == "heat_box_rt") ok = Set2(c->heat_box_rt, v);
      else if (name == "pressure_fixed_point") {
        ok = Set2(c->pressure_fixed_point, v);
        c->has_pressure_fixed_point = 1;
      } else if (name == "initial_sin_n") {
        ok = Set2(c->initial_sin_n, v);
        c->has_initial_sin = 1;
      } else if (name == "box_A" || name == "box_B") {
        // A rigid box holding no cell centre is inert; any other is refused
        // by hg_create once the mesh is known (flag kept conservative).
        if (v.size() >= 2 && (v[0] != 0. || v[1] != 0.)) c->unsupported_flags |= HG_UNSUP_RIGID_BOX;
      }
      if (!ok) return fail(line_no, "vect " + name + " needs two components");
    } else {
      return fail(line_no, "Unknown type '" + type + "'");
    }
  }
  return HG_OK;
}


// ---- 3D (hydro_gpu3d.h) ------------------------------------------------

void hg3_config_default(hg3_config* c) {
  std::memset(c, 0, sizeof(*c));
  hg_config_default(&c->base);
  c->nz = 10;
  c->a_z = 0.;
  c->b_z = 1.;
  c->a1_z = 0.;
  c->b1_z = 1.;  // general.hydroconf: A1 (0, 0, 0), B1 (0, 0, 1)
  for (int s = 0; s < 6; ++s) c->bc_kind[s] = HG_BC_WALL;
}

// The 2D reader takes every key it knows (x and y of vectors); this pass
// adds Nz, the z components and the six wall conditions.
hg_status hg3_config_parse(hg3_config* c, const char* text, char* err, size_t errlen) {
  const hg_status rc = hg_config_parse(&c->base, text, err, errlen);
  if (rc != HG_OK) return rc;
  auto fail = [&](int line_no, const std::string& msg) {
    if (err && errlen) std::snprintf(err, errlen, "line %d: %s", line_no, msg.c_str());
    return HG_ERR_PARSE;
  };
  std::istringstream all(text ? text : "");
  std::string line;
  int line_no = 0;
  while (std::getline(all, line)) {
    ++line_no;
    size_t hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    std::istringstream in(line);
    std::string cmd, type, name;
    if (!(in >> cmd) || cmd != "set") continue;
    if (!(in >> type >> name)) continue;  // refused by the 2D pass
    if (type == "int" && name == "Nz") {
      long v;
      if (!(in >> v)) return fail(line_no, "bad int for Nz");
      c->nz = int32_t(v);
    } else if (type == "vect") {
      std::vector<double> v;
      if (!ReadVect(in, v)) return fail(line_no, "bad vect for " + name);
      const double z = v.size() >= 3 ? v[2] : 0.;
      if (name == "A") c->a_z = z;
      else if (name == "B") c->b_z = z;
      else if (name == "A1") c->a1_z = z;
      else if (name == "B1") c->b1_z = z;
      else if (name == "gravity") c->gravity_z = z;
      else if (name == "force") c->force_z = z;
      else if (name == "initial_velocity") c->initial_velocity_z = z;
      else if (name == "meshvel") c->meshvel_z = z;
    } else if (type == "string" && name.rfind("condition_", 0) == 0) {
      int side = -1;
      if (name == "condition_left") side = HG_SIDE_LEFT;
      else if (name == "condition_right") side = HG_SIDE_RIGHT;
      else if (name == "condition_bottom") side = HG_SIDE_BOTTOM;
      else if (name == "condition_top") side = HG_SIDE_TOP;
      else if (name == "condition_close") side = HG_SIDE_CLOSE;
      else if (name == "condition_far") side = HG_SIDE_FAR;
      if (side < 0) continue;
      std::istringstream cs(RestOfLine(in));
      std::string kind;
      cs >> kind;
      double v3[3] = {0., 0., 0.};
      cs >> v3[0] >> v3[1] >> v3[2];
      if (kind == "wall") c->bc_kind[side] = HG_BC_WALL;
      else if (kind == "inlet") c->bc_kind[side] = HG_BC_INLET;
      else if (kind == "outlet") c->bc_kind[side] = HG_BC_OUTLET;
      else return fail(line_no, "unknown boundary condition '" + kind + "'");
      for (int d = 0; d < 3; ++d) c->bc_velocity[side][d] = kind == "outlet" ? 0. : v3[d];
    }
  }
  return HG_OK;
}

}  // extern "C"
