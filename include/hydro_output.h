/*
 * hydro_output.h -- the hydro2d module's field and scalar output, as a
 * C-ABI over a simulation context (host code in libhydrogpu.so).
 *
 * Replaces the sessions the module builds in hydro::InitOutput
 * (source/hydro2dmpi/hydro2d.hpp:690-926) and drives them as
 * hydro::write_results does (hydro2d.hpp:1623-1652):
 *   - ParaView structured grid: one <filename_field>.<k>.vts per frame plus
 *     the <filename_field>.pvd collection (output_paraview.hpp:34-173);
 *   - plain scalar table <filename_scalar>.dat, one row per scalar frame
 *     (output.hpp:230-261);
 *   - Tecplot ASCII iteration history <filename_iter_history> of the step
 *     with index iter_history_n (hydro2d.hpp:1549-1590,
 *     output_tecplot.hpp:85-118,401-435).
 * Files are byte-compatible with the reference's: values are written as
 * C++ streams write a double by default (6 significant digits, %g).
 * Fields are read from the context (hg_get_field); the writer runs on the
 * host, formatting large arrays on several threads.
 */
#ifndef HYDRO_OUTPUT_H_
#define HYDRO_OUTPUT_H_

#include "hydro_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The module's output entries, in the order it writes them
 * (hydro2d.hpp:711-817); per-phase entries for phases < num_phases. */
typedef enum hg_output_entry {
  HG_OUT_X = 0,                /* nodes: x + meshpos */
  HG_OUT_Y = 1,                /* nodes: y + meshpos */
  HG_OUT_Z = 2,                /* nodes: 0 */
  HG_OUT_VELOCITY_X = 3,       /* cells from here on */
  HG_OUT_VELOCITY_Y = 4,
  HG_OUT_VELOCITY_Z = 5,       /* 0 */
  HG_OUT_PRESSURE = 6,
  HG_OUT_DENSITY = 7,          /* mixture density (unsmoothed) */
  HG_OUT_VISCOSITY = 8,        /* smoothed mixture viscosity */
  HG_OUT_RADIATION = 9,        /* 0: radiation is not run */
  HG_OUT_VOLUME_SOURCE = 10,   /* not available (HG_ERR_UNSUPPORTED) */
  HG_OUT_TEMPERATURE = 11,     /* temperature (temperature_initial without heat_enable) */
  HG_OUT_EXCLUDED = 12,        /* 0: no excluded cells */
  HG_OUT_DIVERGENCE = 13,      /* sum of outward volume fluxes / cell volume */
  HG_OUT_MASS_SOURCE_0 = 14,   /* +i; not available (HG_ERR_UNSUPPORTED) */
  HG_OUT_MASS_FRACTION_0 = 17, /* +i: partial density / mixture density */
  HG_OUT_DENSITY_0 = 20,       /* +i: true density of phase i (density_i) */
  HG_OUT_TARGET_DENSITY_0 = 23,/* +i: density_i (incompressible) */
  HG_OUT_PARTIAL_DENSITY_0 = 26,
  HG_OUT_VOLUME_FRACTION_0 = 29,
  HG_OUT_COUNT = 32
} hg_output_entry;

/* The .hydroconf keys of the output (names as in the reference). */
typedef struct hg_output_config {
  int32_t abi_version;             /* = HG_ABI_VERSION */
  char name[256];                  /* name: experiment name ("exp") */
  char plt_title[256];             /* plt_title ("Template"; "" -> name) */
  char filename_field[512];        /* filename_field ("" -> <name>.field) */
  char filename_scalar[512];       /* filename_scalar ("" -> <name>.scalar) */
  char filename_iter_history[512]; /* filename_iter_history ("" -> <name>.iter_history.plt) */
  char field_output_format[32];    /* "paraview" writes frames; any other value none,
                                      like the reference's disabled "tecplot_binary" */
  int32_t no_output;               /* no_output (0) */
  int32_t no_mesh_output;          /* no_mesh_output (0) */
  int32_t max_frame_index;         /* frames every T/max_frame_index (100) */
  int32_t max_frame_scalar_index;  /* scalar rows every T/max_frame_scalar_index (10000) */
  int32_t output_factor[2];        /* output_factor_x/_y: cells of the output mesh per
                                      cell, per direction (1) */
  double T;                        /* T: end time (1) */
  int32_t iter_history_enable;     /* iter_history_enable (0) */
  int32_t iter_history_n;          /* iter_history_n: index of the step recorded (1) */
  int32_t iter_history_sfixed;     /* iter_history_sfixed: must be 0 */
  int32_t output[HG_OUT_COUNT];    /* output_<entry> flags (examples/general.hydroconf) */
} hg_output_config;

/* Fills the defaults of examples/general.hydroconf (name "exp"). */
void hg_output_config_default(hg_output_config* oc);
/* Applies the output keys of "set <type> <name> <value>" lines; other keys
 * are ignored (hg_config_parse reads those). */
hg_status hg_output_config_parse(hg_output_config* oc, const char* text, char* err,
                                 size_t errlen);

typedef struct hg_output hg_output;

/* Creates the sessions for a context (hydro::InitOutput) and writes what the
 * module's constructor writes: frame 0 at time 0 when max_frame_index > 0,
 * and the first scalar row (hydro2d.hpp:964-971).  Open it right after
 * hg_create (or hg_refresh), as the module does.  Paths are used as given
 * (relative to the working directory).  With no_output the session writes
 * nothing.  HG_ERR_UNSUPPORTED for an entry this build cannot supply. */
hg_status hg_output_open(hg_ctx* ctx, const hg_output_config* oc, hg_output** out);
/* hydro::write_results(force) after a step: a frame when the fluid time
 * passed the next frame time (or force), a scalar row likewise, and the
 * iteration history when the step just finished is number iter_history_n
 * (0-based). */
hg_status hg_output_write_results(hg_output* o, int32_t force);
/* Ends the ParaView collection and closes the files; frees o. */
hg_status hg_output_close(hg_output* o);
/* Message of the last error of o (of the last failed open when o is NULL). */
const char* hg_output_last_error(const hg_output* o);

/* One structured-grid .vts file from host arrays (the frame format of
 * output_paraview.hpp:45-131): nodes xn[nx+1] x yn[ny+1] (tensor grid,
 * node (i,j) -> i + (nx+1)*j), node arrays of (nx+1)*(ny+1) values, cell
 * arrays of nx*ny values (cell (i,j) -> i + nx*j).  HG_ERR_IO when the file
 * cannot be written. */
hg_status hg_vts_write(const char* path, int32_t nx, int32_t ny, const double* xn,
                       const double* yn, int32_t n_node, const char* const* node_names,
                       const double* const* node_values, int32_t n_cell,
                       const char* const* cell_names, const double* const* cell_values);

#ifdef __cplusplus
}
#endif

#endif  /* HYDRO_OUTPUT_H_ */
