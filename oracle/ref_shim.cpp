// ref_shim.cpp -- driver that runs the REFERENCE's own hydro2d module.
//
// TEST INFRASTRUCTURE.  Built by oracle/ref.mk against the reference's
// sources where they lie (never copied into this repository); the binary
// goes to oracle/_ref/.  It applies `set` lines like the reference console
// (console.cpp:276-314 -> BinSearchSet2::read_data_by_key), constructs
// hydro<MeshStructured<double>> directly, calls its step() and dumps raw
// FP64 fields.  Members of the module are private in the reference; the
// driver reads them through member pointers formed in explicit template
// instantiations (where access checks do not apply), which changes no
// reference file and no reference behaviour.
//
// Built with -DHR_DIM=3 (oracle/_ref/hydro_ref3d) it runs the 3D module,
// hydro<geom3d::MeshStructured<double>> (hydro3d.cpp), instead.
//
// usage: hydro_ref <script> <nsteps> <dump-prefix> [dump-every] [hook argument]
//   script: a file of `set <type> <name> <value>` lines
//   writes <prefix>.<step>.bin for each dumped step:
//     header: int32 nx, ny, num_phases, step (3D: nx, ny, nz, num_phases,
//     step); then float64 arrays u[N], v[N] (3D: and w[N]), p[N], flux[F],
//     then alpha_i[N], pd_i[N] per phase,
//     then the temperature T[N]
//   and prints one line per step: "STEP s=<n> iters=<SIMPLE iters> t=<seconds>"
//   dump-every < 0: output mode, no dumps; after each step the module time
//   advances and hydro::write_results runs as in TModule::cycle, so the
//   module's own writers produce their files in the working directory (the
//   run stops early once the time reaches T, like the reference).

#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>
#include <algorithm>
#include <cmath>
#include <array>
#include <exception>
#include <functional>
#include <thread>
#include <mutex>
#include <set>
#include <list>
#include <cstring>

#include "hydro2dmpi/hydro2d.hpp"

#if HR_DIM == 3
using Mesh = geom::geom3d::MeshStructured<double>;
#else
using Mesh = geom::geom2d::MeshStructured<double>;
#endif
using Module = hydro2D_uniform_MPI::hydro<Mesh>;

// Hook run right after the module is constructed (null in hydro_ref; set by
// oracle/ref_gpu_dispatch.cpp in hydro_ref_gpu), and the driver's optional
// fifth argument, which the hook reads.
void (*g_after_construct)(Module&, TExperiment&) = nullptr;
const char* g_extra_arg = nullptr;

// Member access through explicit instantiation.
template <class Tag, typename Tag::type M>
struct Grant {
  friend typename Tag::type Get(Tag) { return M; }
};
#define HR_MEMBER(TAG, TYPE, NAME)                  \
  struct TAG {                                      \
    typedef TYPE Module::*type;                     \
    friend type Get(TAG);                           \
  };                                                \
  template struct Grant<TAG, &Module::NAME>;
HR_MEMBER(TagMesh, Mesh, mesh)
HR_MEMBER(TagFluid, std::shared_ptr<solver::FluidSolver<Mesh>>, fluid_solver)
using AdvPtr = std::shared_ptr<solver::AdvectionSolverMulti<Mesh, geom::FieldFace<double>>>;
HR_MEMBER(TagAdv, AdvPtr, advection_solver)
HR_MEMBER(TagFrac, std::vector<geom::FieldCell<double>>, v_fc_volume_fraction)
HR_MEMBER(TagNp, const size_t, num_phases)
HR_MEMBER(TagHeat, std::shared_ptr<solver::HeatSolver<Mesh>>, heat_solver)
HR_MEMBER(TagSession, std::shared_ptr<output::Session>, session)
HR_MEMBER(TagSessionScalar, std::shared_ptr<output::Session>, session_scalar)
// TModule::increase_time (protected; experiment.hpp:109)
struct TagIncTime {
  typedef void (TModule::*type)();
  friend type Get(TagIncTime);
};
template struct Grant<TagIncTime, &TModule::increase_time>;

static bool ApplySet(TExperiment& ex, const std::string& line) {
  std::stringstream buf(line);
  std::string cmd, type, name;
  buf >> cmd;
  if (cmd != "set") return false;
  buf >> type;
  for (int k = 0; k < Map_number; ++k) {
    if (Map_name[k] == type) {
      buf >> name;
      ex.Map[k]->read_data_by_key(buf, name);
      return true;
    }
  }
  return false;
}

static void Dump(Module& h, const std::string& path, int step) {
  auto& mesh = h.*Get(TagMesh());
  auto& fluid = h.*Get(TagFluid());
  auto& adv = h.*Get(TagAdv());
  auto& frac = h.*Get(TagFrac());
  int nx = int(mesh.GetBlockCells().GetDimensions()[0]);
  int ny = int(mesh.GetBlockCells().GetDimensions()[1]);
  int np = int(h.*Get(TagNp()));
  std::ofstream f(path, std::ios::binary);
#if HR_DIM == 3
  // 3D: header int32 nx, ny, nz, num_phases, step; three velocity components
  int nz = int(mesh.GetBlockCells().GetDimensions()[2]);
  int32_t hdr[5] = {nx, ny, nz, np, step};
  const size_t ncomp = 3;
#else
  int32_t hdr[4] = {nx, ny, np, step};
  const size_t ncomp = 2;
#endif
  f.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
  const auto& vel = fluid->GetVelocity();
  const auto& p = fluid->GetPressure();
  const auto& flux = fluid->GetVolumeFlux();
  std::vector<double> buf;
  for (size_t comp = 0; comp < ncomp; ++comp) {
    buf.clear();
    for (auto c : mesh.Cells()) buf.push_back(vel[c][comp]);
    f.write(reinterpret_cast<const char*>(buf.data()), buf.size() * 8);
  }
  buf.clear();
  for (auto c : mesh.Cells()) buf.push_back(p[c]);
  f.write(reinterpret_cast<const char*>(buf.data()), buf.size() * 8);
  buf.clear();
  for (auto fc : mesh.Faces()) buf.push_back(flux[fc]);
  f.write(reinterpret_cast<const char*>(buf.data()), buf.size() * 8);
  for (int i = 0; i < np; ++i) {
    buf.clear();
    for (auto c : mesh.Cells()) buf.push_back(frac[i][c]);
    f.write(reinterpret_cast<const char*>(buf.data()), buf.size() * 8);
    buf.clear();
    const auto& pd = adv->GetField(i);
    for (auto c : mesh.Cells()) buf.push_back(pd[c]);
    f.write(reinterpret_cast<const char*>(buf.data()), buf.size() * 8);
  }
  // the temperature last (the heat solver always exists; it changes only
  // with heat_enable)
  buf.clear();
  const auto& t = (h.*Get(TagHeat()))->GetTemperature();
  for (auto c : mesh.Cells()) buf.push_back(t[c]);
  f.write(reinterpret_cast<const char*>(buf.data()), buf.size() * 8);
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s <script> <nsteps> <dump-prefix> [dump-every]\n", argv[0]);
    return 2;
  }
  std::string script = argv[1];
  int nsteps = std::atoi(argv[2]);
  std::string prefix = argv[3];
  int every = argc > 4 ? std::atoi(argv[4]) : nsteps;
  try {
    TExperiment ex(nullptr);
    std::ifstream in(script);
    if (!in.good()) {
      std::fprintf(stderr, "cannot open %s\n", script.c_str());
      return 2;
    }
    std::string line;
    while (std::getline(in, line)) {
      auto pos = line.find('#');
      if (pos != std::string::npos) line = line.substr(0, pos);
      if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
      ApplySet(ex, line);
    }
    Module h(&ex);
    g_extra_arg = argc > 5 ? argv[5] : nullptr;
    if (g_after_construct) g_after_construct(h, ex);
    if (every > 0 && prefix != "-") Dump(h, prefix + ".0.bin", 0);
    for (int s = 1; s <= nsteps; ++s) {
      auto t0 = std::chrono::steady_clock::now();
      h.step();
      double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      std::printf("STEP s=%d iters=%d t=%.6f\n", s, int((h.*Get(TagFluid()))->GetIterationCount()),
                  sec);
      std::fflush(stdout);
      if (every > 0 && prefix != "-" && (s % every == 0 || s == nsteps))
        Dump(h, prefix + "." + std::to_string(s) + ".bin", s);
      if (every < 0) {
        // output mode: the rest of TModule::cycle (module.cpp:46-73) --
        // advance the module time and let the module write its results
        (h.*Get(TagIncTime()))();
        bool finished = h.tn >= ex.P_double["T"];
        h.write_results(finished);
        if (finished) break;
      }
    }
    if (every < 0) {
      // close the sessions: the ParaView collection's footer is written by
      // its destructor (output_paraview.hpp:158-160)
      (h.*Get(TagSession())).reset();
      (h.*Get(TagSessionScalar())).reset();
    }
    // TExperiment's destructor expects the console that owns it; this
    // driver has none, so leave without running destructors.
    std::fflush(nullptr);
    std::_Exit(0);
  } catch (const std::string& msg) {
    std::fprintf(stderr, "ERROR: %s\n", msg.c_str());
    return 3;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "ERROR: %s\n", e.what());
    return 3;
  }
  return 0;
}
