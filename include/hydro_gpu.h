/*
 * hydro_gpu.h -- C-ABI of the MI355X hydro2d time-step library (libhydrogpu.so).
 *
 * Plain C types only: no exceptions, no C++ or torch types cross this
 * boundary.  Host buffers passed in are copied, never retained.  Every call
 * returns an hg_status; hg_last_error() gives per-context detail.  A context
 * owns one HIP stream and all of its device buffers and holds no global
 * mutable state, so one context per host thread is safe (the reference runs
 * one experiment per std::thread, source/control/console.cpp:162-178).
 *
 * Field layout is the reference's raw layout (source/hydro2dmpi/mesh.hpp:552-561,
 * 657-713): cell (i,j) -> i + Nx*j; faces are the (Nx+1)*Ny x-faces
 * (i + (Nx+1)*j) followed by the Nx*(Ny+1) y-faces (i + Nx*j).
 */
#ifndef HYDRO_GPU_H_
#define HYDRO_GPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HG_ABI_VERSION 4
#define HG_MAX_PHASES 3

typedef enum hg_status {
  HG_OK = 0,
  HG_ERR_NAN = 1,            /* "NaN pressure"/"NaN field" (fluid.hpp:795-797,1164-1166) */
  HG_ERR_NOT_CONVERGED = 2,  /* solver hit its iteration limit (stats say so) */
  HG_ERR_BAD_ARG = 3,
  HG_ERR_HIP = 4,
  HG_ERR_UNSUPPORTED = 5,    /* a configuration this build does not run */
  HG_ERR_PARSE = 6,          /* .hydroconf text could not be read */
  HG_ERR_IO = 7              /* an output file could not be written */
} hg_status;

/* Linear solver names: the values of linear_solver_velocity /
 * linear_solver_pressure (hydro2d.hpp:195-248).  The first four are the
 * reference's own; the rest are new values of the same keys. */
typedef enum hg_solver {
  HG_SOLVER_LU = 0,           /* "lu": one symmetric Gauss-Seidel sweep (linear.hpp:527-567) */
  HG_SOLVER_LU_RELAXED = 1,   /* "lu_relaxed" (linear.hpp:578-651) */
  HG_SOLVER_GAUSS_SEIDEL = 2, /* "gauss_seidel": lexicographic SOR (linear.hpp:671-716) */
  HG_SOLVER_JACOBI = 3,       /* "jacobi" (linear.hpp:736-783) */
  HG_SOLVER_LU_RB = 4,        /* "lu_rb": red-black ordered symmetric GS sweep (new) */
  HG_SOLVER_CG = 5,           /* "cg": Jacobi-PCG to a relative residual (new) */
  HG_SOLVER_FCG_LEARNED = 6,  /* "fcg_learned": flexible CG + learned preconditioner (new) */
  HG_SOLVER_FCG_MG = 7        /* "fcg_mg": flexible CG + the multigrid V-cycle alone (new) */
} hg_solver;

typedef enum hg_precision {
  HG_PREC_FP64 = 0,           /* every kernel in FP64 */
  HG_PREC_MIXED = 1           /* the step: FP32 per-step face fields, cell gradients,
                                 momentum row coefficients, lu_rb velocity correction and
                                 advection limiter; FP64 state, row constants (residuals),
                                 pressure system, Krylov vectors, dots and corrections.
                                 The pressure solve: FP32 preconditioner (cg: FP32 inner
                                 PCG) under FP64 FCG / iterative refinement */
} hg_precision;

typedef enum hg_bc_kind {
  HG_BC_WALL = 0,             /* "wall vx vy": no-slip wall moving with (vx,vy) */
  HG_BC_INLET = 1,            /* "inlet vx vy": fixed inflow velocity; the phases enter
                                 with the partial densities of the adjacent cells at
                                 construction (hydro2d.hpp:565-583) */
  HG_BC_OUTLET = 2            /* "outlet": the adjacent cell's velocity plus a uniform
                                 normal correction that balances the inflow
                                 (fluid.hpp:542-600) */
} hg_bc_kind;

enum { HG_SIDE_LEFT = 0, HG_SIDE_RIGHT = 1, HG_SIDE_BOTTOM = 2, HG_SIDE_TOP = 3 };

/* Every parameter of the reference's hydro2d module that the time step
 * reads (hydro2d.hpp:267-615,928-972), with the names of the .hydroconf keys.
 * hg_config_default() fills the values of examples/general.hydroconf. */
typedef struct hg_config {
  int32_t abi_version;                 /* = HG_ABI_VERSION */
  int32_t nx, ny;                      /* Nx, Ny */
  double a[2], b[2];                   /* A, B: domain corners */
  double dt;                           /* dt */
  int32_t num_phases;                  /* num_phases (1..3) */
  double density[HG_MAX_PHASES];       /* density_i */
  double viscosity[HG_MAX_PHASES];     /* viscosity_i */
  double gravity[2];                   /* gravity */
  double force[2];                     /* force */
  double sigma;                        /* sigma (surface tension) */
  int32_t bc_kind[4];                  /* condition_left/right/bottom/top kind */
  double bc_velocity[4][2];            /* ... and their velocity */
  /* initial conditions */
  double initial_velocity[2];          /* initial_velocity */
  int32_t has_initial_sin;             /* initial_sin_n present */
  double initial_sin_n[2];             /* initial_sin_n */
  double initial_sin_lambda;           /* initial_sin_lambda */
  double initial_sin_phase;            /* initial_sin_phase */
  double initial_volume_fraction[HG_MAX_PHASES]; /* initial_volume_fraction_i */
  double a1[2], b1[2];                 /* A1, B1: block with phase 1 */
  double a2[2], b2[2];                 /* A2, B2: block with phase 2 */
  double ic[2], ir;                    /* IC, IR: disc with phase 1 */
  double ic2[2], ir2;                  /* IC2, IR2 */
  int32_t initial_volume_fraction_smooth_times;
  int32_t density_smooth_times;
  int32_t viscosity_smooth_times;
  int32_t force_smooth_times;
  /* fluid solver */
  int32_t fluid_enable;                /* fluid_enable */
  int32_t advection_enable;            /* advection_enable */
  int32_t dt_auto;                     /* dt_auto: each step's dt from the volume flux
                                          (hydro2d.hpp:1536-1544) (0) */
  double cfl;                          /* cfl: fluid dt = cfl * min |V / flux| (0.5) */
  double cfl_advection;                /* cfl_advection: advection dt likewise (0.5) */
  double velocity_relaxation_factor;
  double pressure_relaxation_factor;
  double rhie_chow_factor;
  double convergence_tolerance;
  int32_t num_iterations_limit;
  int32_t time_second_order;
  int32_t simpler;                     /* simpler: SIMPLER's second pressure solve per
                                          SIMPLE iteration (fluid.hpp:1059-1151); row
                                          solvers only (0) */
  int32_t force_geometric_average;     /* must be 0 */
  double guess_extrapolation;
  double meshvel[2];
  int32_t has_pressure_fixed_point;    /* pressure_fixed_point present */
  double pressure_fixed_point[2];
  double pressure_fixed_value;
  /* linear solvers */
  int32_t linear_solver_velocity;      /* hg_solver */
  int32_t linear_solver_pressure;      /* hg_solver */
  double lu_relaxed_tolerance;
  int32_t lu_relaxed_num_iters_limit;
  double lu_relaxed_relaxation_factor;
  /* advection */
  double advection_dt_factor;
  int32_t tvd_split;
  double sharp;
  /* heat: a passive temperature advected by the step's flux and diffused
   * (hydro2d.hpp:652-688,1607-1612; heat.hpp:18-94) */
  int32_t heat_enable;                 /* heat_enable (0) */
  int32_t linear_solver_heat;          /* linear_solver_heat: lu, lu_rb, gauss_seidel, jacobi,
                                          lu_relaxed (lu) */
  double heat_box_lb[2], heat_box_rt[2]; /* heat_box_lb/_rt: boundary faces whose centre
                                          lies in this box hold heat_box_temperature; the
                                          others have zero normal derivative ((0,0)) */
  double conductivity[HG_MAX_PHASES];  /* conductivity_i (1) */
  double temperature_initial;          /* temperature_initial (0) */
  double heat_box_temperature;         /* heat_box_temperature (0) */
  double heat_relaxation_factor;       /* heat_relaxation_factor (1) */
  int32_t time_second_order_heat;      /* time_second_order_heat (1) */
  /* new keys of the GPU path (no reference counterpart) */
  double pressure_tolerance_gpu;       /* relative residual for cg / fcg_learned (1e-8) */
  int32_t pressure_max_iters_gpu;      /* iteration cap for cg / fcg_learned */
  int32_t precision_gpu;               /* hg_precision */
  int32_t learned_levels_gpu;          /* levels of the learned preconditioner (4) */
  int32_t learned_channels_gpu;        /* channels at the finest level (16) */
  int32_t pressure_warm_start_gpu;     /* 1: a step's first cg / fcg solve starts from the
                                          previous step's correction (1) */
  int32_t fcg_negate_precond_gpu;      /* check knob of the FCG fallback: 1 makes flexible CG
                                          apply minus its preconditioner, so (r, z) < 0 on
                                          the first iteration and the solve continues with
                                          FP64 Jacobi-PCG to the same tolerance (0) */
  /* output side of the module state */
  int32_t meshvel_output;              /* meshvel_output: the output mesh moves with meshvel
                                          (hydro2d.hpp:1526-1528) (1) */
  /* Bits of hg_unsupported: features the script switched on that this path
   * does not run (hg_create refuses them with HG_ERR_UNSUPPORTED). */
  int32_t unsupported_flags;
} hg_config;

typedef enum hg_unsupported {
  HG_UNSUP_DT_AUTO = 1 << 0,        /* (no longer set: dt_auto runs) */
  HG_UNSUP_COMPRESSIBLE = 1 << 1,   /* compressible_enable 1 */
  HG_UNSUP_HEAT = 1 << 2,           /* (no longer set: the heat solver runs) */
  HG_UNSUP_RADIATION = 1 << 3,      /* radiation_enable 1 */
  HG_UNSUP_CHEMISTRY = 1 << 4,      /* chem_intensity != 0 */
  HG_UNSUP_ANTIDIFFUSION = 1 << 5,  /* antidiffusion_factor set */
  HG_UNSUP_SETTLING = 1 << 6,       /* enable_settling_i 1 */
  HG_UNSUP_ADVECTION = 1 << 7,      /* advection_solver other than tvd */
  HG_UNSUP_RIGID_BOX = 1 << 8,      /* box_A/box_B holding cells */
  HG_UNSUP_CONDITION = 1 << 9,      /* (no longer set: inlets and outlets run) */
  HG_UNSUP_IMAGE_INIT = 1 << 10,    /* img_init / imgu_init / imgv_init */
  HG_UNSUP_OTHER = 1 << 11          /* initial_pois, deforming_velocity, meshvel_auto */
} hg_unsupported;

typedef enum hg_field {
  HG_FIELD_VELOCITY_X = 0,     /* cells, time_curr layer */
  HG_FIELD_VELOCITY_Y = 1,
  HG_FIELD_PRESSURE = 2,
  HG_FIELD_VOLUME_FLUX = 3,    /* faces, time_curr layer */
  HG_FIELD_PARTIAL_DENSITY_0 = 4,   /* +i for phase i */
  HG_FIELD_VOLUME_FRACTION_0 = 7,   /* +i for phase i (read only) */
  HG_FIELD_DENSITY = 10,            /* unsmoothed mixture density (read only) */
  HG_FIELD_DENSITY_SMOOTH = 11,     /* (read only) */
  HG_FIELD_VISCOSITY_SMOOTH = 12,   /* (read only) */
  HG_FIELD_FORCE_X = 13,            /* (read only) */
  HG_FIELD_FORCE_Y = 14,            /* (read only) */
  HG_FIELD_VOLUME_FLUX_ITER = 15,   /* faces, iter_curr layer (advection input) */
  /* The last SIMPLE pressure-correction system in face form (read only):
   * k_f = A/(d a_f) on inner faces, 0 on walls, and the cell right-hand
   * side b with A p' = b (fluid.hpp:946-1027). */
  HG_FIELD_PRESSURE_FACE_COEF = 16, /* faces */
  HG_FIELD_PRESSURE_RHS = 17,       /* cells */
  HG_FIELD_TEMPERATURE = 18,        /* cells, time_curr layer (heat_enable) */
  HG_FIELD_COUNT = 19
} hg_field;

typedef struct hg_iter_stats {
  int32_t iteration;            /* SIMPLE iteration count after this call ("s") */
  double convergence_indicator; /* "Rs" (fluid.hpp:174-179) */
  int32_t velocity_solve_iters[2];
  int32_t pressure_solve_iters; /* sweeps / CG iterations of the pressure solve */
  double pressure_solve_diff;   /* final max|corr| (SOR/Jacobi) or relative residual (CG) */
  int32_t pressure_converged;   /* 1 if the pressure solver met its tolerance */
} hg_iter_stats;

typedef struct hg_step_stats {
  int32_t simple_iterations;
  double convergence_indicator;     /* Rs after the last SIMPLE iteration */
  int32_t advection_substeps;
  int64_t pressure_solve_iters_total;
  double time;                      /* fluid time after the step */
  double dt;                        /* the step's time step (dt_auto: chosen by the step) */
  double pressure_solve_diff_max;   /* the largest final pressure_solve_diff of the step's
                                       solves (cg / fcg: true FP64 relative residual) */
  int32_t pressure_converged;       /* 1 if every pressure solve of the step met its
                                       tolerance */
} hg_step_stats;

typedef struct hg_stats {           /* CalcStat (hydro2d.hpp:1432-1529), per phase */
  double volume[HG_MAX_PHASES];
  double mass[HG_MAX_PHASES];
  double pd_min[HG_MAX_PHASES];
  double pd_max[HG_MAX_PHASES];
  double cx[HG_MAX_PHASES], cy[HG_MAX_PHASES];   /* + the output offset meshpos */
  double vx[HG_MAX_PHASES], vy[HG_MAX_PHASES];
  /* volume and mass through the boundary, in and out, integrated in time:
   * every property update adds dt * sum over boundary faces of the flux
   * into (out of) the domain times the cell's volume fraction
   * (hydro2d.hpp:1474-1508); zero on walls at rest */
  double volume_in[HG_MAX_PHASES], volume_out[HG_MAX_PHASES];
  double mass_in[HG_MAX_PHASES], mass_out[HG_MAX_PHASES];
} hg_stats;

typedef struct hg_progress {        /* what the module's output reads besides fields */
  double time;                      /* fluid time */
  double module_time;               /* the module's "t" after the console's time advance
                                       (module.cpp:74-81): 0, then t + the dt in force
                                       before the step, so with dt_auto it trails the fluid
                                       time by one step's change of dt */
  double dt, dt_advection;          /* time steps in force (dt_auto changes them) */
  int64_t steps;                    /* steps finished (FinishStep calls) */
  int32_t iteration;                /* "s": SIMPLE iteration count of the last iteration
                                       made (hydro2d.hpp:1580-1581); 0 before any */
  double convergence_indicator;     /* FluidSimple::GetConvergenceIndicator: 1 before the
                                       first SIMPLE iteration of a step, else the last Rs
                                       (fluid.hpp:174-179,1169-1171) */
  double meshpos[2];                /* offset of the output mesh (hydro2d.hpp:1526-1528) */
} hg_progress;

/* ---- configuration (the .hydroconf interface) ------------------------- */

/* Fills the defaults of examples/general.hydroconf. */
void hg_config_default(hg_config* cfg);

/* Applies "set <int|double|string|bool|vect> <name> <value>" lines
 * (console.cpp:276-314) to cfg.  Unknown keys are ignored like the reference
 * stores them unused; a known key with an unreadable value is HG_ERR_PARSE. */
hg_status hg_config_parse(hg_config* cfg, const char* text, char* err, size_t errlen);

/* Maps a solver name ("lu", "gauss_seidel", "cg", ...) to hg_solver; -1 if unknown. */
int32_t hg_solver_from_name(const char* name);

/* ---- context lifecycle (replaces hydro<Mesh>::hydro, hydro2d.hpp:928-972) */

typedef struct hg_ctx hg_ctx;

hg_status hg_create(const hg_config* cfg, hg_ctx** out);
void hg_destroy(hg_ctx* ctx);
/* The context's last error; with NULL, that of the last failed hg_create
 * on this thread. */
const char* hg_last_error(const hg_ctx* ctx);
int32_t hg_num_cells(const hg_ctx* ctx);
int32_t hg_num_faces(const hg_ctx* ctx);

/* Field transfer (host <-> HBM).  set_field writes both time layers of the
 * field; call hg_refresh() afterwards to recompute what the constructor
 * derives from it (initial volume flux, fluid properties, stats). */
hg_status hg_set_field(hg_ctx* ctx, int32_t field, const double* host, size_t n);
hg_status hg_get_field(hg_ctx* ctx, int32_t field, double* host, size_t n);
hg_status hg_refresh(hg_ctx* ctx);
/* The configuration the context was created with. */
hg_status hg_get_config(const hg_ctx* ctx, hg_config* cfg);
/* Time, step and iteration counters, convergence indicator, mesh offset. */
hg_status hg_get_progress(hg_ctx* ctx, hg_progress* out);
/* Convergence indicators Rs of the SIMPLE iterations of the current step (of
 * the last one once it finished), in order: *n receives their number, and
 * the first min(*n, cap) go to rs (rs may be NULL with cap 0).  The
 * reference logs them per iteration (hydro2d.hpp:1583-1590). */
hg_status hg_get_iteration_history(hg_ctx* ctx, double* rs, int32_t cap, int32_t* n);

/* ---- stepping at the reference's granularity --------------------------- */

/* FluidSimple::StartStep / MakeIteration / IsConverged / FinishStep
 * (fluid.hpp:793-1169, solver.hpp:733-736). */
hg_status hg_fluid_start_step(hg_ctx* ctx);
hg_status hg_fluid_iterate(hg_ctx* ctx, hg_iter_stats* stats);
int32_t hg_fluid_is_converged(const hg_ctx* ctx);
hg_status hg_fluid_finish_step(hg_ctx* ctx);
/* One AdvectionSolverMultiExplicit sub-step (advection.hpp:417-545). */
hg_status hg_advect_substep(hg_ctx* ctx);
/* hydro::UpdateFluidProperties (hydro2d.hpp:1404-1430) and the running part
 * of hydro::CalcStat that follows it in the step (the boundary in/out
 * integrals and meshpos, hydro2d.hpp:1474-1528). */
hg_status hg_update_properties(hg_ctx* ctx);
/* hydro::CalcStat reductions (hydro2d.hpp:1432-1529) of the current state. */
hg_status hg_calc_stats(hg_ctx* ctx, hg_stats* stats);
/* hydro::step (hydro2d.hpp:1531-1621): with dt_auto the step's dt first
 * (FluidSimple::GetAutoTimeStep, fluid.hpp:1191-1206), then the SIMPLE loop,
 * advection sub-steps, property update, stats.  Returns once the SIMPLE loop
 * has finished on the GPU (its NaN check and *stats are settled); the
 * advection, property and stats kernels may still be running on the
 * context's stream.  Every call that reads device state waits for them. */
hg_status hg_step(hg_ctx* ctx, hg_step_stats* stats);
/* Wall-clock of the last hg_step split by the reference's timer regions
 * (step.fluid, step.advection, step.fluid_properties), in ms, HIP events;
 * waits for the step's last kernels.  Zeros before the first hg_step. */
hg_status hg_last_step_timing(const hg_ctx* ctx, double* fluid_ms, double* advection_ms,
                              double* properties_ms);

/* Average device time (ms) of one launch of a step kernel on this context's
 * current state, launched `reps` times back to back between two HIP events
 * on its stream (no reference counterpart; for throughput figures).  The
 * launches write scratch outputs only, so the state is unchanged.
 * kernel: "KAdvStep" (one advection sub-step of all phase fields),
 * "KMomentum" (momentum rows of one SIMPLE iteration). */
hg_status hg_time_kernel(hg_ctx* ctx, const char* kernel, int32_t reps, double* avg_ms);

/* ---- many cases at once (replaces the console's experiment threads,
 * console.cpp:162-178 and :923, which run each experiment's step loop on its
 * own std::thread, at most max_threads_number at a time) ---------------------
 * Steps every context `steps` times.  Up to `threads` host threads (0: one
 * per context) each take whole contexts from a shared queue; every context
 * runs on its own HIP stream, so the cases overlap on the GPU.  Results are
 * those of stepping each context alone (the contexts share nothing).
 * stats[i] (may be NULL) receives the last step's stats of context i;
 * *wall_ms the wall time of the whole batch.  Returns the first error met
 * (by context order); the other contexts still run.  Every context may
 * appear only once in ctxs (HG_ERR_BAD_ARG otherwise). */
hg_status hg_batch_run(hg_ctx* const* ctxs, int32_t n, int32_t steps, int32_t threads,
                       hg_step_stats* stats, double* wall_ms);

/* ---- solve level (replaces LinearSolver<Scal,IdxCell,Expr>::Solve,
 * linear.hpp:302-310) ------------------------------------------------------
 * A 5-point system on an nx*ny grid: row r means
 *   s[r]*x[r-nx] + w[r]*x[r-1] + c[r]*x[r] + e[r]*x[r+1] + n[r]*x[r+nx] + k[r] = 0
 * (terms outside the grid must be 0).  x is overwritten with the solution
 * from a zero initial guess, as every reference solver starts from 0. */
typedef struct hg_solve_opts {
  int32_t solver;            /* hg_solver */
  double tolerance;          /* SOR/Jacobi: max|corr|; CG: relative residual */
  int32_t max_iters;         /* SOR/Jacobi: num_iters_limit; CG: iteration cap */
  double relaxation_factor;  /* SOR/Jacobi omega */
  int32_t precision;         /* hg_precision */
} hg_solve_opts;

typedef struct hg_solve_stats {
  int32_t iterations;
  double final_diff;         /* max|corr| (SOR/Jacobi) */
  double rel_residual;       /* ||b - A x||_2 / ||b||_2 in FP64, b = -k */
  int32_t converged;
  double ms;                 /* device time of the solve (HIP events) */
} hg_solve_stats;

hg_status hg_stencil_solve(int32_t nx, int32_t ny, const double* s, const double* w,
                           const double* c, const double* e, const double* n,
                           const double* k, double* x, const hg_solve_opts* opts,
                           hg_solve_stats* stats);

/* ---- variable-coefficient pressure Poisson (div(1/rho grad p) = div u*) --
 * The SIMPLE pressure-correction system (fluid.hpp:946-1027) in face form:
 * kx[(nx+1)*ny], ky[nx*(ny+1)] are the face coefficients A/(d*a_f) (>0,
 * zero on walls), rhs[nx*ny] the cell right-hand side.  Operator:
 *   (A p)_c = sum_f k_f (p_c - p_nb)  (symmetric positive semi-definite).
 * A pressure-solve object keeps the operator and the learned preconditioner
 * resident in HBM so repeated solves do not re-upload. */
typedef struct hg_poisson hg_poisson;

hg_status hg_poisson_create(int32_t nx, int32_t ny, const double* kx, const double* ky,
                            hg_poisson** out);
void hg_poisson_destroy(hg_poisson* ps);
/* Solves A p = rhs (rhs projected to zero mean) from p = 0 to the relative
 * residual opts->tolerance (true FP64 residual).  opts->solver is
 * HG_SOLVER_CG (Jacobi-PCG), HG_SOLVER_FCG_LEARNED or HG_SOLVER_FCG_MG.
 * opts->precision: for cg, HG_PREC_MIXED runs FP32 inner solves under FP64
 * iterative refinement; for fcg_mg, HG_PREC_MIXED preconditions with the
 * FP32 V-cycle (the step's) and HG_PREC_FP64 with the same V-cycle on FP64
 * levels (FCG itself is FP64 either way). */
hg_status hg_poisson_solve(hg_poisson* ps, const double* rhs, double* p,
                           const hg_solve_opts* opts, hg_solve_stats* stats);

/* Average device time (ms) of one launch of a pressure-solve kernel on this
 * object's operator, launched `reps` times back to back between two HIP
 * events on its stream (no reference counterpart; for throughput figures).
 * kernel: "KDown<double>" / "KUp<double, true>" (finest multigrid level),
 * "KCoarseCycle" (the one-workgroup coarse V-cycle).  The data are whatever
 * the level buffers hold; call after a solve. */
hg_status hg_poisson_time_kernel(hg_poisson* ps, const char* kernel, int32_t reps,
                                 double* avg_ms);

/* ---- learned preconditioner (no reference counterpart: the reference's
 * pressure solvers are the stationary sweeps of linear.hpp) ----------------
 * HG_SOLVER_FCG_LEARNED preconditions flexible CG with
 *   M(r) = V(r) + N(D^-1 (r - A V(r)))
 * (V: multigrid V-cycle, N: a 4-level BF16 U-Net, 16..64 channels) once the
 * net is trained, and with V alone before.  Training pairs come from the
 * solver on the operator in use: random multi-scale right-hand sides, a
 * tight reference solve, then the stationary V-cycle iteration's residuals
 * and remaining errors.  Requires nx, ny multiples of 8 (>= 16). */
typedef struct hg_train_opts {
  int32_t rhs_count;   /* right-hand sides (4) */
  int32_t mg_iters;    /* V-cycle iterations per right-hand side = pairs each (12) */
  int32_t steps;       /* Adam steps, one pair each (300) */
  float lr;            /* peak learning rate, cosine decay to lr/10 (1e-3) */
  uint64_t seed;       /* (1) */
  int32_t flags;       /* HG_TRAIN_* bits (HG_TRAIN_TIED) */
} hg_train_opts;

#define HG_TRAIN_LINEAR 1      /* network without activations: M stays a linear operator */
#define HG_TRAIN_STANDALONE 2  /* no V-cycle: M(r) = N(D^-1 r) (S N S when tied); judged
                                  against Jacobi-PCG */
#define HG_TRAIN_TIED 4        /* symmetric linear net N = sum_l 4^l F_l^T F_l: M = V + S N S,
                                  S = D^-1/2, symmetric positive definite by construction */

typedef struct hg_train_stats {
  int32_t pairs;
  int32_t steps;
  double loss_first;   /* mean relative energy-norm error (best output multiple),
                          first 10% of steps */
  double loss_last;    /* ... last 10% */
  double ms;           /* wall time of pair generation, training and acceptance */
  /* Acceptance solve on a held-out right-hand side to a 1e-8 true FP64
   * relative residual: FCG with the learned preconditioner against the
   * method without it (FCG + V-cycle; Jacobi-PCG for HG_TRAIN_STANDALONE).
   * The solvers use the net only when it took fewer iterations and less
 * device time. */
  int32_t accepted;
  int32_t val_iters;
  int32_t val_iters_baseline;
  double val_ms;
  double val_ms_baseline;
} hg_train_stats;

void hg_train_opts_default(hg_train_opts* opts);
/* Trains the preconditioner of a pressure-solve object on its operator. */
hg_status hg_poisson_train(hg_poisson* ps, const hg_train_opts* opts, hg_train_stats* stats);
/* Trains the preconditioner of a simulation on the operator of its last
 * pressure solve (linear_solver_pressure "fcg_learned"; at least one
 * iteration done). */
hg_status hg_train_pressure_net(hg_ctx* ctx, const hg_train_opts* opts, hg_train_stats* stats);

/* ---- the network alone (no reference counterpart), for checks and
 * throughput figures.  Arrays are host arrays of nx*ny floats (cell
 * i + nx*j); weights are packed per layer as W[tap][ci][co], tap = 3*ky+kx,
 * layers E0..E3, D3..D0. */
typedef struct hg_net hg_net;
hg_status hg_net_create(int32_t nx, int32_t ny, hg_net** out);
void hg_net_destroy(hg_net* net);
int64_t hg_net_num_weights(const hg_net* net);
/* HG_TRAIN_LINEAR / HG_TRAIN_TIED bits: activations off / tied decoder. */
hg_status hg_net_configure(hg_net* net, int32_t flags);
hg_status hg_net_set_weights(hg_net* net, const float* w);
hg_status hg_net_get_weights(hg_net* net, float* w);
/* y = N(x) */
hg_status hg_net_forward(hg_net* net, const float* x, float* y);
/* gw = d/dw sum(gy * N(x)) */
hg_status hg_net_gradient(hg_net* net, const float* x, const float* gy, float* gw);
/* One Adam step on mean((N(x) - t)^2); *loss = that mean before the step. */
hg_status hg_net_train_step(hg_net* net, const float* x, const float* t, float lr, double* loss);
/* Device time per forward pass and per training step on resident synthetic
 * input (HIP events over `reps` launches each), and the FLOPs of each. */
hg_status hg_net_time(hg_net* net, int32_t reps, double* forward_ms, double* train_ms,
                      double* forward_flops, double* train_flops);

#ifdef __cplusplus
}
#endif

#endif  /* HYDRO_GPU_H_ */
