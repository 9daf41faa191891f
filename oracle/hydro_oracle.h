/*
 * hydro_oracle.h -- CPU oracle of the hydro2d time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C++ restatement of the
 * reference algorithm (divfree/hydro, source/hydro2dmpi/{hydro2d,fluid,
 * conv_diff,advection,solver,linear,mesh2d}.hpp) used as the parity checker
 * and as the CPU baseline.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product library
 * (hydro_amd/libhydrogpu.so) never links or calls it.
 *
 * The API mirrors include/hydro_gpu.h so a test can drive both sides with
 * the same hg_config and the same call sequence.
 */
#ifndef HYDRO_ORACLE_H_
#define HYDRO_ORACLE_H_

#include "../include/hydro_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

int or_create(const hg_config* cfg, or_ctx** out, char* err, size_t errlen);
void or_destroy(or_ctx* ctx);
int or_num_cells(const or_ctx* ctx);
int or_num_faces(const or_ctx* ctx);
int or_set_field(or_ctx* ctx, int field, const double* host, size_t n);
int or_get_field(or_ctx* ctx, int field, double* host, size_t n);
int or_refresh(or_ctx* ctx);

int or_fluid_start_step(or_ctx* ctx);
int or_fluid_iterate(or_ctx* ctx, hg_iter_stats* stats);
int or_fluid_is_converged(const or_ctx* ctx);
int or_fluid_finish_step(or_ctx* ctx);
int or_advect_substep(or_ctx* ctx);
int or_update_properties(or_ctx* ctx);
int or_calc_stats(or_ctx* ctx, hg_stats* stats);
int or_step(or_ctx* ctx, hg_step_stats* stats);

/* Accumulated wall time (seconds) of a reference timer region, by name
 * ("step", "step.fluid", "fluid.6.pressure-solve", ...); -1 if unknown. */
double or_timer(const or_ctx* ctx, const char* region);

/* Last momentum and pressure systems (for solver-level tests):
 * which = 0: velocity rows (shared coefficients), constants of component 0
 *         in k, of component 1 in k2 (k2 may be NULL);
 * which = 1: pressure-correction rows. */
int or_get_system(or_ctx* ctx, int which, double* s, double* w, double* c, double* e,
                  double* n, double* k, double* k2);

/* Solver level: same contract as hg_stencil_solve. */
int or_stencil_solve(int nx, int ny, const double* s, const double* w, const double* c,
                     const double* e, const double* n, const double* k, double* x,
                     const hg_solve_opts* opts, hg_solve_stats* stats);

/* Variable-coefficient Poisson in face form (same contract as hg_poisson_solve,
 * Jacobi-PCG only). */
int or_poisson_solve(int nx, int ny, const double* kx, const double* ky, const double* rhs,
                     double* p, const hg_solve_opts* opts, hg_solve_stats* stats);

/* Reference geometry of the uniform mesh (mesh.hpp:722-736, mesh2d.hpp:215-245),
 * for the geometry tests: cell centers/volumes, face centers/areas/surfaces. */
// CPU baselines of the pressure solve to a relative residual: method 0 serial
// lexicographic SOR (the reference's algorithm, residual checked every
// check_every sweeps), method 1 OpenMP Jacobi-PCG.
int or_poisson_baseline(int nx, int ny, const double* kx, const double* ky, const double* rhs,
                        int method, double omega, double tol, int max_iters, int check_every,
                        double* p, double* seconds, int* iters, double* rel_res);
int or_geometry(const or_ctx* ctx, double* cell_cx, double* cell_cy, double* cell_vol,
                double* face_cx, double* face_cy, double* face_area, double* face_sx,
                double* face_sy);

/* BDF coefficients (solver.hpp:815-861) for the tests. */
void or_bdf_coeffs(double dt, int second_order, double out[3]);

#ifdef __cplusplus
}
#endif

#endif
