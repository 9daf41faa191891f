// output.cpp -- the module's field and scalar output (host code).
//
// The sessions of hydro::InitOutput (hydro2d.hpp:690-926) driven as
// hydro::write_results (hydro2d.hpp:1623-1652), writing the reference's
// formats: ParaView .vts frames and the .pvd collection
// (output_paraview.hpp:34-173), the plain scalar table
// (output.hpp:230-261) and the Tecplot ASCII iteration history
// (output_tecplot.hpp:85-118,401-435).  The text layout of those files is
// an interface (other tools read them), so it is reproduced byte for byte;
// numbers are written as a C++ stream writes a double by default, which is
// printf's %g (precision 6).  Fields come from the context through the
// C-ABI; long arrays are formatted on several host threads.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/hydro_output.h"
#include "host_geometry.h"

namespace {

struct OutError : std::runtime_error {
  hg_status code;
  OutError(hg_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// ---- number formatting -------------------------------------------------------

// A double as `std::ostream << double` writes it with the default flags and
// precision: printf's "%g".
inline int FormatG(char* buf, size_t n, double v) { return std::snprintf(buf, n, "%g", v); }

std::string Fmt(double v) {
  char b[40];
  FormatG(b, sizeof b, v);
  return b;
}

// Appends "<v> " for each value.  Arrays of more than one chunk are
// formatted chunk by chunk on up to 16 threads and joined in order.
constexpr long kChunk = 4096;

void AppendValues(std::string& out, const double* v, long n) {
  auto one = [v, n](long c, std::string& s) {
    const long b = c * kChunk, e = std::min(n, b + kChunk);
    s.reserve(size_t(e - b) * 13);
    char buf[48];
    for (long q = b; q < e; ++q) {
      int k = FormatG(buf, sizeof buf - 1, v[q]);
      buf[k] = ' ';
      s.append(buf, size_t(k) + 1);
    }
  };
  const long nchunks = (n + kChunk - 1) / kChunk;
  if (nchunks <= 1) {
    if (n > 0) one(0, out);
    return;
  }
  std::vector<std::string> parts(static_cast<size_t>(nchunks));
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int nt = int(std::min<long>(nchunks, long(std::min(hw, 16u))));
  std::atomic<long> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&]() {
      for (long c; (c = next.fetch_add(1)) < nchunks;) one(c, parts[size_t(c)]);
    });
  for (auto& t : pool) t.join();
  for (auto& p : parts) out += p;
}

// ---- the .vts frame (output_paraview.hpp:45-131) -----------------------------

using Array = std::pair<std::string, const double*>;

void DataArrayHeader(std::string& s, const std::string& name, int comps) {
  s += "        <DataArray Name=\"" + name + "\" NumberOfComponents=\"" + std::to_string(comps) +
       "\" type=\"Float32\" format=\"ascii\">\n";
}
const char kDataArrayFooter[] = "        </DataArray>\n";

void WriteVts(const std::string& path, int nx, int ny, const double* xn, const double* yn,
              const std::vector<Array>& nodes, const std::vector<Array>& cells) {
  std::ofstream f(path, std::ios::binary);
  if (!f) throw OutError(HG_ERR_IO, "cannot write " + path);
  const long nn = long(nx + 1) * (ny + 1), nc = long(nx) * ny;
  std::string s;
  auto flush = [&]() {
    f.write(s.data(), std::streamsize(s.size()));
    s.clear();
  };
  const std::string ext =
      "0 " + std::to_string(nx) + " 0 " + std::to_string(ny) + " 0 0";
  s += "<?xml version=\"1.0\"?>\n";
  s += "<VTKFile type=\"StructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
  s += "  <StructuredGrid WholeExtent=\"" + ext + "\">\n";
  s += "    <Piece Extent=\"" + ext + "\">\n";
  s += "      <PointData>\n";
  for (const auto& a : nodes) {
    DataArrayHeader(s, a.first, 1);
    AppendValues(s, a.second, nn);
    s += "\n";
    s += kDataArrayFooter;
    flush();
  }
  s += "      </PointData>\n";
  s += "      <CellData>\n";
  for (const auto& a : cells) {
    DataArrayHeader(s, a.first, 1);
    AppendValues(s, a.second, nc);
    s += "\n";
    s += kDataArrayFooter;
    flush();
  }
  s += "      </CellData>\n";
  s += "      <Points>\n";
  DataArrayHeader(s, "mesh", 3);
  {
    // node coordinates with a zero third component; the reference ends the
    // list without a line break
    std::vector<double> pts(size_t(nn) * 3);
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        const size_t q = size_t(i + long(nx + 1) * j) * 3;
        pts[q] = xn[i];
        pts[q + 1] = yn[j];
        pts[q + 2] = 0.;
      }
    AppendValues(s, pts.data(), nn * 3);
  }
  s += kDataArrayFooter;
  s += "      </Points>\n";
  s += "    </Piece>\n";
  s += "  </StructuredGrid>\n";
  s += "</VTKFile>\n";
  flush();
  f.close();
  if (!f) throw OutError(HG_ERR_IO, "cannot write " + path);
}

// ---- configuration -----------------------------------------------------------

const char* const kEntryNames[14] = {
    "x",        "y",         "z",         "velocity_x",    "velocity_y",
    "velocity_z", "pressure", "density",  "viscosity",     "radiation",
    "volume_source", "temperature", "excluded", "divergence"};
const char* const kPhaseEntryNames[6] = {"mass_source_",    "mass_fraction_",
                                         "density_",        "target_density_",
                                         "partial_density_", "volume_fraction_"};

std::string EntryName(int e) {
  if (e < HG_OUT_MASS_SOURCE_0) return kEntryNames[e];
  const int k = e - HG_OUT_MASS_SOURCE_0;
  return std::string(kPhaseEntryNames[k / HG_MAX_PHASES]) + std::to_string(k % HG_MAX_PHASES);
}

int EntryFromName(const std::string& n) {
  for (int e = 0; e < HG_OUT_COUNT; ++e)
    if (n == EntryName(e)) return e;
  return -1;
}

void SetString(char* dst, size_t cap, const std::string& v) {
  std::snprintf(dst, cap, "%s", v.c_str());
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

thread_local std::string g_open_error;

}  // namespace

// ---- the session ---------------------------------------------------------------

struct hg_output {
  hg_ctx* ctx = nullptr;
  hg_config cfg{};
  hg_output_config oc{};
  std::string err;
  bool off = false;     // no_output: nothing is written
  bool frames = false;  // a ParaView session exists
  std::string field_stem, title, history_path;
  std::ofstream pvd, dat;
  long frame = 0;  // index of the next .vts (SessionParaviewStructured::timestep_)
  double last_frame_time = 0., last_frame_scalar_time = 0.;
  long history_step = -1;
  int onx = 0, ony = 0;
  std::vector<double> oxn, oyn;  // nodes of the output mesh
  hg::HostMesh mesh;             // the simulation mesh (cell volumes)

  std::vector<double> Get(int field, long n) {
    std::vector<double> v(static_cast<size_t>(n));
    hg_status rc = hg_get_field(ctx, field, v.data(), size_t(n));
    if (rc != HG_OK) throw OutError(rc, std::string("hg_get_field: ") + hg_last_error(ctx));
    return v;
  }
  hg_progress Progress() {
    hg_progress p{};
    hg_status rc = hg_get_progress(ctx, &p);
    if (rc != HG_OK) throw OutError(rc, std::string("hg_get_progress: ") + hg_last_error(ctx));
    return p;
  }

  // SessionParaviewStructured::Write (output_paraview.hpp:163-168) of the
  // entries of hydro2d.hpp:711-817 evaluated on the output mesh.
  void WriteFrame(double time) {
    const std::string file = field_stem + "." + std::to_string(frame) + ".vts";
    pvd << "    <DataSet timestep=\"" << Fmt(time) << "\" group=\"\" part=\"0\" file=\"" << file
        << "\"/>\n";
    pvd.flush();
    const int nx = cfg.nx, ny = cfg.ny, np = cfg.num_phases;
    const long nc = long(nx) * ny, nf = long(nx + 1) * ny + long(nx) * (ny + 1);
    const int fx = oc.output_factor[0], fy = oc.output_factor[1];
    const long onn = long(onx + 1) * (ony + 1), onc = long(onx) * ony;
    const hg_progress pr = Progress();
    std::vector<std::vector<double>> store;  // owns the arrays `nodes` and `cells` point into
    store.reserve(HG_OUT_COUNT);
    std::vector<Array> nodes, cells;
    auto on = [&](int e) { return oc.output[e] != 0; };
    // node entries
    for (int e = HG_OUT_X; e <= HG_OUT_Z; ++e) {
      if (!on(e)) continue;
      std::vector<double> v(static_cast<size_t>(onn));
      for (int j = 0; j <= ony; ++j)
        for (int i = 0; i <= onx; ++i) {
          double x = 0.;
          if (e == HG_OUT_X) x = oxn[i] + pr.meshpos[0];
          if (e == HG_OUT_Y) x = oyn[j] + pr.meshpos[1];
          v[size_t(i + long(onx + 1) * j)] = x;
        }
      store.push_back(std::move(v));
      nodes.emplace_back(EntryName(e), store.back().data());
    }
    // cell entries on the simulation mesh, then spread over the output mesh
    std::vector<double> rho;
    for (int e = HG_OUT_VELOCITY_X; e < HG_OUT_COUNT; ++e) {
      if (!on(e)) continue;
      int phase = -1, group = -1;
      if (e >= HG_OUT_MASS_SOURCE_0) {
        phase = (e - HG_OUT_MASS_SOURCE_0) % HG_MAX_PHASES;
        group = (e - HG_OUT_MASS_SOURCE_0) / HG_MAX_PHASES;
        if (phase >= np) continue;  // the module lists phases < num_phases only
      }
      std::vector<double> v;
      if (e == HG_OUT_VELOCITY_X) v = Get(HG_FIELD_VELOCITY_X, nc);
      else if (e == HG_OUT_VELOCITY_Y) v = Get(HG_FIELD_VELOCITY_Y, nc);
      else if (e == HG_OUT_PRESSURE) v = Get(HG_FIELD_PRESSURE, nc);
      else if (e == HG_OUT_DENSITY) v = Get(HG_FIELD_DENSITY, nc);
      else if (e == HG_OUT_VISCOSITY) v = Get(HG_FIELD_VISCOSITY_SMOOTH, nc);
      else if (e == HG_OUT_VELOCITY_Z || e == HG_OUT_RADIATION || e == HG_OUT_EXCLUDED) {
        v.assign(size_t(nc), 0.);
      } else if (e == HG_OUT_TEMPERATURE) {
        // the heat solver exists in every run; without heat_enable its
        // temperature stays temperature_initial (hydro2d.hpp:652-688)
        if (cfg.heat_enable) v = Get(HG_FIELD_TEMPERATURE, nc);
        else v.assign(size_t(nc), cfg.temperature_initial);
      } else if (e == HG_OUT_DIVERGENCE) {
        // sum over the cell's faces (left, right, bottom, top; outward
        // factors -1, 1, -1, 1) of the volume flux, over the cell volume
        // (hydro2d.hpp:753-761, mesh2d.hpp:118-129,281-285)
        const std::vector<double> fl = Get(HG_FIELD_VOLUME_FLUX, nf);
        const long nfx = long(nx + 1) * ny;
        v.resize(size_t(nc));
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i) {
            double sum = 0.;
            sum += fl[size_t(i + long(nx + 1) * j)] * -1.;
            sum += fl[size_t(i + 1 + long(nx + 1) * j)] * 1.;
            sum += fl[size_t(nfx + i + long(nx) * j)] * -1.;
            sum += fl[size_t(nfx + i + long(nx) * (j + 1))] * 1.;
            v[size_t(i + long(nx) * j)] = sum / mesh.Volume(i, j);
          }
      } else if (group == 1) {  // mass fraction: partial density / mixture density
        if (rho.empty()) rho = Get(HG_FIELD_DENSITY, nc);
        v = Get(HG_FIELD_PARTIAL_DENSITY_0 + phase, nc);
        for (long c = 0; c < nc; ++c) v[size_t(c)] = v[size_t(c)] / rho[size_t(c)];
      } else if (group == 2 || group == 3) {  // true and target density: density_i
        v.assign(size_t(nc), cfg.density[phase]);
      } else if (group == 4) {
        v = Get(HG_FIELD_PARTIAL_DENSITY_0 + phase, nc);
      } else if (group == 5) {
        v = Get(HG_FIELD_VOLUME_FRACTION_0 + phase, nc);
      } else {
        throw OutError(HG_ERR_UNSUPPORTED, "output_" + EntryName(e) + " is not available");
      }
      if (fx != 1 || fy != 1) {
        std::vector<double> w(static_cast<size_t>(onc));
        for (int j = 0; j < ony; ++j)
          for (int i = 0; i < onx; ++i)
            w[size_t(i + long(onx) * j)] = v[size_t(i / fx + long(nx) * (j / fy))];
        v.swap(w);
      }
      store.push_back(std::move(v));
      cells.emplace_back(EntryName(e), store.back().data());
    }
    WriteVts(file, onx, ony, oxn.data(), oyn.data(), nodes, cells);
    ++frame;
  }

  // SessionPlainScalar::Write (output.hpp:244-258) of the entries of
  // hydro2d.hpp:825-893.
  void WriteScalarRow() {
    const hg_progress pr = Progress();
    hg_stats st{};
    hg_status rc = hg_calc_stats(ctx, &st);
    if (rc != HG_OK) throw OutError(rc, std::string("hg_calc_stats: ") + hg_last_error(ctx));
    std::string s;
    auto put = [&s](double v) {
      s += Fmt(v);
      s += ' ';
    };
    put(pr.module_time);  // P_double["t"], the module's clock
    put(double(pr.iteration));
    put(pr.convergence_indicator);
    const int np = cfg.num_phases;
    for (int i = 0; i < np; ++i) {
      put(st.mass[i]);
      put(st.mass_in[i]);
      put(st.mass_out[i]);
    }
    for (int i = 0; i < np; ++i) {
      put(st.volume[i]);
      put(st.volume_in[i]);
      put(st.volume_out[i]);
    }
    for (int i = 0; i < np; ++i) {
      put(st.pd_min[i]);
      put(st.pd_max[i]);
    }
    for (int i = 0; i < np; ++i) {
      put(st.cx[i]);
      put(st.cy[i]);
      put(st.vx[i]);
      put(st.vy[i]);
    }
    s += '\n';
    dat << s;
    dat.flush();
    if (!dat) throw OutError(HG_ERR_IO, "cannot write the scalar table");
  }

  // The iteration history of one step (hydro2d.hpp:1549-1590): header of
  // SessionTecplotAscii / SessionTecplotAsciiScalar, then "iter Rs " rows.
  void WriteHistory() {
    int32_t n = 0;
    hg_status rc = hg_get_iteration_history(ctx, nullptr, 0, &n);
    std::vector<double> rs(size_t(std::max(n, 1)));
    if (rc == HG_OK) rc = hg_get_iteration_history(ctx, rs.data(), n, &n);
    if (rc != HG_OK) throw OutError(rc, std::string("history: ") + hg_last_error(ctx));
    std::ofstream f(history_path, std::ios::binary);
    if (!f) throw OutError(HG_ERR_IO, "cannot write " + history_path);
    std::string s;
    s += "Title=\"" + title + "\"\n";
    s += "Variables=\"iter\" \"iter_diff_velocity\" \n\n";
    s += "Zone ZoneType=ORDERED DataPacking=POINT\n";
    s += "T=\"" + title + "\"\n";
    for (int32_t k = 0; k < n; ++k) {
      s += Fmt(double(k + 1));
      s += ' ';
      s += Fmt(rs[size_t(k)]);
      s += " \n";
    }
    f << s;
    f.close();
    if (!f) throw OutError(HG_ERR_IO, "cannot write " + history_path);
  }

  void Open() {
    if (hg_get_config(ctx, &cfg) != HG_OK) throw OutError(HG_ERR_BAD_ARG, "bad context");
    if (oc.no_output) {  // InitOutput and write_results return at once
      off = true;
      return;
    }
    if (oc.iter_history_enable && oc.iter_history_sfixed > 0)
      throw OutError(HG_ERR_UNSUPPORTED,
                     "iter_history_sfixed > 0 (iterating past convergence) is not supported");
    if (oc.output[HG_OUT_VOLUME_SOURCE])
      throw OutError(HG_ERR_UNSUPPORTED, "output_volume_source is not available");
    for (int i = 0; i < cfg.num_phases; ++i)
      if (oc.output[HG_OUT_MASS_SOURCE_0 + i])
        throw OutError(HG_ERR_UNSUPPORTED,
                       "output_" + EntryName(HG_OUT_MASS_SOURCE_0 + i) + " is not available");
    if (oc.output_factor[0] < 1 || oc.output_factor[1] < 1)
      throw OutError(HG_ERR_BAD_ARG, "output_factor_x/_y must be >= 1");
    // output mesh: InitUniformMesh on the same domain (hydro2d.hpp:695-708)
    onx = cfg.nx * oc.output_factor[0];
    ony = cfg.ny * oc.output_factor[1];
    oxn.resize(size_t(onx) + 1);
    oyn.resize(size_t(ony) + 1);
    const double dx = cfg.b[0] - cfg.a[0], dy = cfg.b[1] - cfg.a[1];
    for (int i = 0; i <= onx; ++i) oxn[size_t(i)] = cfg.a[0] + (double(i) / double(onx)) * dx;
    for (int j = 0; j <= ony; ++j) oyn[size_t(j)] = cfg.a[1] + (double(j) / double(ony)) * dy;
    mesh.Init(cfg.nx, cfg.ny, cfg.a, cfg.b);
    const std::string name = oc.name;
    title = oc.plt_title[0] ? std::string(oc.plt_title) : name;
    field_stem = oc.filename_field[0] ? std::string(oc.filename_field) : name + ".field";
    const std::string scalar =
        oc.filename_scalar[0] ? std::string(oc.filename_scalar) : name + ".scalar";
    history_path = oc.filename_iter_history[0] ? std::string(oc.filename_iter_history)
                                               : name + ".iter_history.plt";
    if (std::string(oc.field_output_format) == "paraview") {
      frames = true;
      pvd.open(field_stem + ".pvd", std::ios::binary);
      if (!pvd) throw OutError(HG_ERR_IO, "cannot write " + field_stem + ".pvd");
      pvd << "<?xml version=\"1.0\"?>\n"
          << "<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
          << "  <Collection>\n";
      pvd.flush();
    }
    dat.open(scalar + ".dat", std::ios::binary);
    if (!dat) throw OutError(HG_ERR_IO, "cannot write " + scalar + ".dat");
    {
      // header: entry names (hydro2d.hpp:830-893, output.hpp:234-240)
      std::string s = "time num_iters iter_diff_velocity ";
      const int np = cfg.num_phases;
      auto ph = [](const char* b, int i) { return std::string(b) + std::to_string(i) + " "; };
      for (int i = 0; i < np; ++i) s += ph("mass_", i) + ph("mass_in_", i) + ph("mass_out_", i);
      for (int i = 0; i < np; ++i)
        s += ph("volume_", i) + ph("volume_in_", i) + ph("volume_out_", i);
      for (int i = 0; i < np; ++i) s += ph("pd_min_", i) + ph("pd_max_", i);
      for (int i = 0; i < np; ++i) s += ph("cx_", i) + ph("cy_", i) + ph("vx_", i) + ph("vy_", i);
      dat << s << "\n";
    }
    // what the constructor writes (hydro2d.hpp:964-971)
    if (oc.max_frame_index > 0 && frames) WriteFrame(0.);
    WriteScalarRow();
    last_frame_time = 0.;
    last_frame_scalar_time = 0.;
  }

  // hydro::write_results (hydro2d.hpp:1623-1652).
  void WriteResults(bool force) {
    if (off) return;
    const hg_progress pr = Progress();
    if (oc.iter_history_enable && pr.steps >= 1 && pr.steps - 1 == oc.iter_history_n &&
        history_step != pr.steps) {
      WriteHistory();
      history_step = long(pr.steps);
    }
    const double time = pr.time;  // fluid_solver->GetTime()
    const double total = oc.T;
    // the frame counts are size_t in the reference: total / count in double
    const double frame_duration = total / double(size_t(long(oc.max_frame_index)));
    if (force || (!oc.no_mesh_output && time >= last_frame_time + frame_duration)) {
      last_frame_time += frame_duration;
      if (frames) WriteFrame(time);
    }
    const double scalar_duration = total / double(size_t(long(oc.max_frame_scalar_index)));
    if (force || time >= last_frame_scalar_time + scalar_duration) {
      last_frame_scalar_time = time;
      WriteScalarRow();
    }
  }

  void Close() {
    if (pvd.is_open()) {
      pvd << "  </Collection>\n</VTKFile>\n";
      pvd.close();
      if (!pvd) throw OutError(HG_ERR_IO, "cannot write the .pvd collection");
    }
    if (dat.is_open()) dat.close();
  }
};

namespace {

hg_status OutGuard(std::string& err, const std::function<void()>& fn) {
  try {
    fn();
    return HG_OK;
  } catch (const OutError& e) {
    err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    err = e.what();
    return HG_ERR_IO;
  }
}

}  // namespace

extern "C" {

void hg_output_config_default(hg_output_config* oc) {
  std::memset(oc, 0, sizeof(*oc));
  oc->abi_version = HG_ABI_VERSION;
  SetString(oc->name, sizeof oc->name, "exp");
  SetString(oc->plt_title, sizeof oc->plt_title, "Template");
  SetString(oc->field_output_format, sizeof oc->field_output_format, "paraview");
  oc->max_frame_index = 100;
  oc->max_frame_scalar_index = 10000;
  oc->output_factor[0] = oc->output_factor[1] = 1;
  oc->T = 1.;
  oc->iter_history_n = 1;
  for (int e : {int(HG_OUT_X), int(HG_OUT_Y), int(HG_OUT_VELOCITY_X), int(HG_OUT_VELOCITY_Y),
                int(HG_OUT_PRESSURE), int(HG_OUT_VOLUME_FRACTION_0),
                int(HG_OUT_VOLUME_FRACTION_0) + 1})
    oc->output[e] = 1;
}

hg_status hg_output_config_parse(hg_output_config* oc, const char* text, char* err,
                                 size_t errlen) {
  if (!oc) return HG_ERR_BAD_ARG;
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
    if (!(in >> type >> name)) return fail(line_no, "incomplete set");
    if (type == "int" || type == "bool") {
      long v;
      if (!(in >> v)) return fail(line_no, "bad " + type + " for " + name);
      const int32_t iv = int32_t(v);
      if (name == "no_output") oc->no_output = iv;
      else if (name == "no_mesh_output") oc->no_mesh_output = iv;
      else if (name == "max_frame_index") oc->max_frame_index = iv;
      else if (name == "max_frame_scalar_index") oc->max_frame_scalar_index = iv;
      else if (name == "output_factor_x") oc->output_factor[0] = iv;
      else if (name == "output_factor_y") oc->output_factor[1] = iv;
      else if (name == "iter_history_enable") oc->iter_history_enable = iv;
      else if (name == "iter_history_n") oc->iter_history_n = iv;
      else if (name == "iter_history_sfixed") oc->iter_history_sfixed = iv;
      else if (name.rfind("output_", 0) == 0) {
        const int e = EntryFromName(name.substr(7));
        if (e >= 0) oc->output[e] = iv;
      }
    } else if (type == "double") {
      double v;
      if (!(in >> v)) return fail(line_no, "bad double for " + name);
      if (name == "T") oc->T = v;
    } else if (type == "string") {
      const std::string v = RestOfLine(in);
      if (name == "name") SetString(oc->name, sizeof oc->name, v);
      else if (name == "plt_title") SetString(oc->plt_title, sizeof oc->plt_title, v);
      else if (name == "filename_field") SetString(oc->filename_field, sizeof oc->filename_field, v);
      else if (name == "filename_scalar")
        SetString(oc->filename_scalar, sizeof oc->filename_scalar, v);
      else if (name == "filename_iter_history")
        SetString(oc->filename_iter_history, sizeof oc->filename_iter_history, v);
      else if (name == "field_output_format")
        SetString(oc->field_output_format, sizeof oc->field_output_format, v);
    }
  }
  return HG_OK;
}

hg_status hg_output_open(hg_ctx* ctx, const hg_output_config* oc, hg_output** out) {
  if (!out) return HG_ERR_BAD_ARG;
  *out = nullptr;
  if (!ctx || !oc) {
    g_open_error = "null context or configuration";
    return HG_ERR_BAD_ARG;
  }
  if (oc->abi_version != HG_ABI_VERSION) {
    g_open_error = "hg_output_config from another ABI version";
    return HG_ERR_BAD_ARG;
  }
  std::unique_ptr<hg_output> o(new hg_output());
  o->ctx = ctx;
  o->oc = *oc;
  hg_status rc = OutGuard(g_open_error, [&]() { o->Open(); });
  if (rc != HG_OK) return rc;
  *out = o.release();
  return HG_OK;
}

hg_status hg_output_write_results(hg_output* o, int32_t force) {
  if (!o) return HG_ERR_BAD_ARG;
  return OutGuard(o->err, [&]() { o->WriteResults(force != 0); });
}

hg_status hg_output_close(hg_output* o) {
  if (!o) return HG_ERR_BAD_ARG;
  std::string err;
  hg_status rc = OutGuard(err, [&]() { o->Close(); });
  if (rc != HG_OK) g_open_error = err;
  delete o;
  return rc;
}

const char* hg_output_last_error(const hg_output* o) {
  return o ? o->err.c_str() : g_open_error.c_str();
}

hg_status hg_vts_write(const char* path, int32_t nx, int32_t ny, const double* xn,
                       const double* yn, int32_t n_node, const char* const* node_names,
                       const double* const* node_values, int32_t n_cell,
                       const char* const* cell_names, const double* const* cell_values) {
  if (!path || nx < 1 || ny < 1 || !xn || !yn || n_node < 0 || n_cell < 0 ||
      (n_node > 0 && (!node_names || !node_values)) ||
      (n_cell > 0 && (!cell_names || !cell_values)))
    return HG_ERR_BAD_ARG;
  std::vector<Array> nodes, cells;
  for (int32_t k = 0; k < n_node; ++k) {
    if (!node_names[k] || !node_values[k]) return HG_ERR_BAD_ARG;
    nodes.emplace_back(node_names[k], node_values[k]);
  }
  for (int32_t k = 0; k < n_cell; ++k) {
    if (!cell_names[k] || !cell_values[k]) return HG_ERR_BAD_ARG;
    cells.emplace_back(cell_names[k], cell_values[k]);
  }
  std::string err;
  return OutGuard(err, [&]() { WriteVts(path, nx, ny, xn, yn, nodes, cells); });
}

}  // extern "C"
