/* hydro_gpu3d.h -- C-ABI of the 3D time step (hydro3d) on MI355X.
 *
 * Replaces hydro<geom::geom3d::MeshStructured<double>>::step, the module the
 * reference registers as "hydro3d" (hydro3d.cpp:17-20) on the structured 3D
 * mesh (mesh3d.hpp:20-464): the same SIMPLE loop, advection and property
 * update as the 2D step, with six faces per cell and 7-point rows.
 *
 * Cells are indexed i + Nx (j + Ny k); faces in the reference's order
 * (mesh.hpp:641-720): the (Nx+1) Ny Nz x-faces (i + (Nx+1)(j + Ny k)), then
 * the Nx (Ny+1) Nz y-faces, then the Nx Ny (Nz+1) z-faces.
 *
 * Scope of this path: one or two phases (the second set in the box A1-B1),
 * walls on all six sides (condition_left/right/bottom/top/close/far
 * "wall vx vy vz"), gravity and a uniform force, the row solvers jacobi,
 * gauss_seidel and lu (the oracle) -- on the GPU jacobi (bitwise the
 * reference's), cg (FP64 Jacobi-PCG to pressure_tolerance_gpu) and, for the
 * pressure, fcg_mg (FP64 flexible CG with an FP32 multigrid V-cycle, to the
 * same tolerance; Nx, Ny, Nz must halve while even to at most 4096 cells).
 * Everything else (heat, SIMPLER, dt_auto, inlets and outlets, mesh
 * velocity, surface tension, sharpening, split TVD) is refused by
 * hg3_create with HG_ERR_UNSUPPORTED.
 */
#ifndef HYDRO_GPU3D_H
#define HYDRO_GPU3D_H

#include "hydro_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { HG_SIDE_CLOSE = 4, HG_SIDE_FAR = 5 }; /* z = A_z and z = B_z */

/* Field ids of the 3D context: the 2D ids, plus the z components. */
enum {
  HG_FIELD_VELOCITY_Z = 32, /* cells, time_curr layer */
  HG_FIELD_FORCE_Z = 33     /* (read only) */
};

typedef struct hg3_config {
  hg_config base;                /* every key the 2D path reads (x and y of vectors) */
  int32_t nz;                    /* Nz */
  double a_z, b_z;               /* z of A, B */
  double a1_z, b1_z;             /* z of A1, B1 */
  double gravity_z, force_z;     /* z of gravity, force */
  double initial_velocity_z;     /* z of initial_velocity */
  double meshvel_z;              /* z of meshvel (must be 0) */
  int32_t bc_kind[6];            /* left right bottom top close far (HG_BC_*) */
  double bc_velocity[6][3];      /* their wall velocities */
} hg3_config;

/* Defaults of general.hydroconf with Nz 10 and B (1, 1, 1). */
void hg3_config_default(hg3_config* cfg);
/* Applies `set` lines (the 2D grammar; vectors take three components). */
hg_status hg3_config_parse(hg3_config* cfg, const char* text, char* err, size_t errlen);

typedef struct hg3_ctx hg3_ctx;

/* hydro::hydro (hydro2d.hpp:194-250) on the 3D mesh: fields, conditions and
 * solvers from the configuration; refuses what this path does not run. */
hg_status hg3_create(const hg3_config* cfg, hg3_ctx** out);
void hg3_destroy(hg3_ctx* ctx);
/* The message of the last failed call on ctx (or of the last hg3_create). */
const char* hg3_last_error(const hg3_ctx* ctx);
/* hydro::step (hydro2d.hpp:1531-1621). */
hg_status hg3_step(hg3_ctx* ctx, hg_step_stats* stats);
/* Fields of the time_curr layer, by HG_FIELD_* id; n must equal the field's
 * size (cells or faces).  set accepts velocity, pressure, flux and partial
 * densities and resets the previous time layer to the same values. */
hg_status hg3_get_field(hg3_ctx* ctx, int32_t field, double* out, size_t n);
hg_status hg3_set_field(hg3_ctx* ctx, int32_t field, const double* in, size_t n);

#ifdef __cplusplus
}
#endif

#endif /* HYDRO_GPU3D_H */
