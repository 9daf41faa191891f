"""Algorithmic HBM bytes per launch of every kernel on the time-step path.

Counting rule: each array a kernel must read or write between two global
synchronisations counts once, at its element size (FP64 8 B, FP32 4 B,
BF16 2 B), whatever the kernel actually re-reads through the caches.  N is
the number of cells (nx*ny), F the number of faces ((nx+1)*ny + nx*(ny+1)),
Fx/Fy the x-/y-faces.  Level-l multigrid arrays have N_l = N / 4^l cells.

The same table gives `roofline.achieved` in bench.py (bytes / measured
duration) and the per-kernel bandwidth column of the profile summaries.
"""

F64, F32 = 8, 4


def sizes(nx, ny):
    n = nx * ny
    fx, fy = (nx + 1) * ny, nx * (ny + 1)
    return n, fx + fy, fx, fy


def _b(nx, ny, cells64=0, faces64=0, cells32=0, faces32=0, coarse32=0):
    n, f, _, _ = sizes(nx, ny)
    return (cells64 * n + faces64 * f) * F64 + (cells32 * n + faces32 * f + coarse32 * n) * F32


# name -> (bytes(nx, ny), what is counted)
KERNELS = {
    # ---- SIMPLE iteration (step_kernels.hip) ----
    "KForceFaces": (lambda nx, ny: _b(nx, ny, cells64=3, faces64=3),
                    "read f_x, f_y, mu; write face f_x, f_y, mu"),
    "KRestoredForce": (lambda nx, ny: _b(nx, ny, cells64=2, faces64=1),
                       "read face f_x (x-faces), f_y (y-faces); write cell restored force x, y"),
    "KFacesDeriv0Pair": (lambda nx, ny: _b(nx, ny, cells64=2, faces64=2),
                         "read 2 cell fields; write 2 face fields"),
    "KCellGrads<false>": (lambda nx, ny: _b(nx, ny, cells64=9),
                   "read p, u0, u1; write 3 gradients x 2 components"),
    "KMomentum": (lambda nx, ny: _b(nx, ny, cells64=23, faces64=2),
                  "read u0, u1, 4 velocity gradients, grad p (2), restored force (2), rho, "
                  "u at t-1 and t (4), face mu, volume flux; write 5 row coefficients, "
                  "2 constants and a_P"),
    "KPressureRhs": (lambda nx, ny: _b(nx, ny, cells64=1, faces64=1),
                     "read flux*; write the pressure rows' constant"),
    "KLuRbBlack": (lambda nx, ny: _b(nx, ny, cells64=8),
                   "read 5 coefficients, 2 constants; write the black half of 2 iterates"),
    "KLuRbRed": (lambda nx, ny: _b(nx, ny, cells64=13),
                 "read 5 coefficients, 2 constants, 2 black iterates, 2 bases; write 2 results"),
    "KRhieChow": (lambda nx, ny: _b(nx, ny, cells64=6, faces64=4),
                  "read u0, u1, grad p (2), p, a_P, face restored force and force; "
                  "write flux* and flux coefficient"),
    "KPressureRows": (lambda nx, ny: _b(nx, ny, cells64=6, faces64=2),
                      "read flux*, flux coefficient; write 5 row coefficients + constant"),
    "KCorrectCells": (lambda nx, ny: _b(nx, ny, cells64=10),
                      "read p', a_P, p_prev, u0, u1, u_prev (2); write p, u0, u1"),
    "KCorrectFaces": (lambda nx, ny: _b(nx, ny, cells64=1, faces64=3),
                      "read flux*, flux coefficient, p'; write flux"),
    # ---- the same stages in the mixed path: FP32 intermediates ----
    "KForceFaces mixed": (lambda nx, ny: _b(nx, ny, cells64=3, faces32=3),
                          "read f_x, f_y, mu (FP64); write face f_x, f_y, mu (FP32)"),
    "KRestoredForce mixed": (lambda nx, ny: _b(nx, ny, faces32=1, cells32=2),
                             "read face f_x, f_y (FP32); write restored force x, y (FP32)"),
    "KFacesDeriv0Pair mixed": (lambda nx, ny: _b(nx, ny, cells32=2, faces32=2),
                               "read 2 cell fields; write 2 face fields (FP32)"),
    "KCellGrads<false> mixed": (lambda nx, ny: _b(nx, ny, cells64=3, cells32=6),
                                "read p, u0, u1 (FP64); write 3 gradients x 2 components (FP32)"),
    "KMomentum mixed": (lambda nx, ny: _b(nx, ny, cells64=8, cells32=15, faces64=1, faces32=1),
                        "read u0, u1, rho, u at t-1 and t (4), volume flux (FP64) and 4 "
                        "velocity gradients, grad p (2), restored force (2), face mu (FP32); "
                        "write 5 row coefficients and 2 constants (FP32) and a_P (FP64)"),
    "KLuRbBlack mixed": (lambda nx, ny: _b(nx, ny, cells32=8),
                         "read 5 coefficients, 2 constants; write 2 black corrections (FP32)"),
    "KLuRbRed mixed": (lambda nx, ny: _b(nx, ny, cells32=9, cells64=4),
                       "read 5 coefficients, 2 constants, 2 black corrections (FP32), 2 bases "
                       "(FP64); write 2 results (FP64)"),
    "KRhieChow mixed": (lambda nx, ny: _b(nx, ny, cells64=4, cells32=2, faces64=2, faces32=2),
                        "read u0, u1, p, a_P (FP64), grad p (FP32), face restored force and "
                        "force (FP32); write flux* and flux coefficient (FP64)"),
    "KCorrectFaces mixed": (lambda nx, ny: _b(nx, ny, cells64=1, faces64=3, faces32=1),
                            "read flux*, flux coefficient, p'; write flux (FP64) and its FP32 "
                            "copy"),
    # ---- pressure-solve set-up and closing check (cg.hip) ----
    "KSetOperator": (lambda nx, ny: _b(nx, ny, faces64=2, faces32=1, cells64=1, cells32=1),
                     "read the flux coefficients; write k_f (FP64 and FP32), D^-1 (FP64, FP32)"),
    "KSumPart": (lambda nx, ny: _b(nx, ny, cells64=1), "read b (partial sums of the mean)"),
    "KStart": (lambda nx, ny: _b(nx, ny, cells64=4), "read b; write projected b, r, x = 0"),
    "KResidual": (lambda nx, ny: _b(nx, ny, cells64=3, faces64=1), "read b, x, k_x, k_y; write r"),
    # ---- advection and properties ----
    "KAdvStep<2, false>": (lambda nx, ny: _b(nx, ny, cells64=4, faces64=1),
                    "two phase fields: read both fields and the volume flux; write both fields"),
    "KAdvStep<2, false> mixed": (lambda nx, ny: _b(nx, ny, cells64=4, faces32=1),
                                 "two phase fields (FP64): read both and the FP32 volume flux; "
                                 "write both"),
    "KAdvGrad": (lambda nx, ny: _b(nx, ny, cells64=3), "read field; write gradient (2)"),
    "KAdvUpdate": (lambda nx, ny: _b(nx, ny, cells64=4, faces64=1),
                   "read field, gradient (2), volume flux; write field"),
    "KProperties": (lambda nx, ny: _b(nx, ny, cells64=8),
                    "2 phases: read 2 partial densities; write 2 fractions, rho, mu, force (2)"),
    "KSmooth": (lambda nx, ny: _b(nx, ny, cells64=2), "read field; write smoothed field"),
    "KSmooth2": (lambda nx, ny: _b(nx, ny, cells64=2),
                 "two smoothing passes: read field; write twice-smoothed field"),
    "KExtrapolate": (lambda nx, ny: _b(nx, ny, cells64=6, faces64=2),
                     "guess_extrapolation 0: read time_curr u_x, u_y, p, flux; write "
                     "iter_curr (time_prev only where time_curr is zero or non-finite)"),
    # ---- FP64 Jacobi-PCG iteration (cg.hip) ----
    "KPcgSpmv<double>": (lambda nx, ny: _b(nx, ny, cells64=4, faces64=1),
                         "read z, p_old, k_x, k_y; write p, q"),
    "KPcgUpdate<double>": (lambda nx, ny: _b(nx, ny, cells64=8),
                           "read p, q, D^-1, x, r; write x, r, z"),
    # ---- FCG iteration with the FP32 multigrid preconditioner ----
    "KFcgCheck": (lambda nx, ny: 0, "one workgroup folds partial sums (preconditioners "
                  "without the fused stop test)"),
    "KFcgDir": (lambda nx, ny: _b(nx, ny, cells64=2, cells32=1),
                "read z (FP32), q_old, r (preconditioners without the fused dots)"),
    "KFcgSpmv": (lambda nx, ny: _b(nx, ny, cells64=3, faces64=1, cells32=1),
                 "read z (FP32), p_old, k_x, k_y; write p, q"),
    "KFcgUpdate": (lambda nx, ny: _b(nx, ny, cells64=6), "read p, q, x, r; write x, r"),
    # level 0 of the V-cycle (FP32 level arrays, FP64 residual input)
    "KDown<double>": (lambda nx, ny: _b(nx, ny, cells64=1, cells32=2, faces32=1, coarse32=0.25),
                      "read r (FP64), D^-1, k_x, k_y; write x2; write the coarse residual (N/4)"),
    "KUp<double, true>": (lambda nx, ny: _b(nx, ny, cells64=1, cells32=3, faces32=1.25,
                                            coarse32=0.25),
                          "read x2, r (FP64), D^-1, k_x, k_y, coarse correction and coarse k; "
                          "write the level output"),
    # in an FCG iteration it also reads q_old for the direction dots
    "KUp<double, true> + FCG dots": (lambda nx, ny: _b(nx, ny, cells64=2, cells32=3,
                                                       faces32=1.25, coarse32=0.25),
                                     "as KUp<double, true>, plus q_old"),
    "KDown<double, 16, true>": (lambda nx, ny: _b(nx, ny, cells64=6, cells32=2, faces32=1,
                                                  coarse32=0.25),
                                "KDown<double> with the previous FCG update: plus read q, "
                                "p, x; write x and the new r"),
}


def kernel_bytes(name, nx, ny):
    """Algorithmic bytes of one launch of `name` on an nx x ny grid (for the
    multigrid kernels: on that level)."""
    fn, _ = KERNELS[name]
    return float(fn(nx, ny))


def cg_iteration_bytes(nx, ny):
    """FP64 Jacobi-PCG iteration: SpMV + update (16 words per cell counting
    k_x, k_y as one word each)."""
    return kernel_bytes("KPcgSpmv<double>", nx, ny) + kernel_bytes("KPcgUpdate<double>", nx, ny)


def _vcycle_bytes(nx, ny, fcg_dots=False):
    """One V-cycle: the finest level's fused kernels plus the FP32 levels
    below it (each level quarters the cells; down-leg reads r, D^-1, k and
    writes x2 and the coarse residual; up-leg reads x2, r, D^-1, k and the
    coarse correction and writes the output).  fcg_dots: the V-cycle of a
    flexible CG iteration: the finest up-leg also reads q_old (the direction
    dots) and the finest down-leg carries the previous iteration's update."""
    up = "KUp<double, true> + FCG dots" if fcg_dots else "KUp<double, true>"
    down = "KDown<double, 16, true>" if fcg_dots else "KDown<double>"
    b = kernel_bytes(down, nx, ny) + kernel_bytes(up, nx, ny)
    cx, cy = nx // 2, ny // 2
    while cx * cy > 2048 and cx % 2 == 0 and cy % 2 == 0:
        b += _b(cx, cy, cells32=3, faces32=1, coarse32=0.25)      # down
        b += _b(cx, cy, cells32=4, faces32=1.25, coarse32=0.25)   # up
        cx, cy = cx // 2, cy // 2
    return b


def headline_step_bytes(nx, ny, simple_iters, fcg_iters, substeps=10, mixed=True):
    """Algorithmic bytes of one step of the headline path (two phases,
    lu_rb velocity, FCG + V-cycle pressure; mixed: the FP32 intermediates of
    the mixed path): the sum over the step's kernels, launches per step times
    bytes per launch."""
    n, f, _, _ = sizes(nx, ny)
    w = F64
    m = " mixed" if mixed else ""
    start = w * 3 * n + kernel_bytes("KExtrapolate", nx, ny) \
        + kernel_bytes("KForceFaces" + m, nx, ny) + kernel_bytes("KRestoredForce" + m, nx, ny) \
        + kernel_bytes("KFacesDeriv0Pair" + m, nx, ny)             # NaN check, guesses, faces
    per_iter = sum(kernel_bytes(k + (m if k != "KPressureRhs" and k != "KCorrectCells" else ""),
                                nx, ny) for k in (
        "KCellGrads<false>", "KMomentum", "KLuRbBlack", "KLuRbRed", "KRhieChow", "KPressureRhs",
        "KCorrectCells", "KCorrectFaces"))
    # operator set-up (the multigrid's finest level shares the FP32 arrays;
    # coarse levels are negligible), FCG start and the closing true residual
    per_iter += sum(kernel_bytes(k, nx, ny) for k in (
        "KSetOperator", "KSumPart", "KStart", "KResidual"))
    # FCG iteration: stop test and direction dots inside the V-cycle kernels
    # (the update of each iteration runs in the next V-cycle's first kernel)
    per_fcg = kernel_bytes("KFcgSpmv", nx, ny) + _vcycle_bytes(nx, ny, fcg_dots=True)
    finish = w * 3 * n                                             # NaN check
    advection = substeps * kernel_bytes("KAdvStep<2, false>" + m, nx, ny)
    # density and viscosity, two smoothing passes each in one launch
    properties = kernel_bytes("KProperties", nx, ny) + 2 * kernel_bytes("KSmooth2", nx, ny)
    return start + simple_iters * per_iter + fcg_iters * per_fcg + finish + advection + properties
