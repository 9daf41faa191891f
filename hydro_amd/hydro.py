"""Host-side mirror of the reference's hydro2d module over the C-ABI.

`Hydro` follows hydro<Mesh> (source/hydro2dmpi/hydro2d.hpp): it is built
from .hydroconf text (the same `set` lines the reference console reads),
steps with `step()` (hydro::step, hydro2d.hpp:1531-1621) and exposes the
FluidSimple protocol (StartStep / MakeIteration / IsConverged / FinishStep,
fluid.hpp:793-1169).  Errors mirror the reference's: a NaN raises
`HydroNaN("NaN pressure")` like the `throw std::string("NaN pressure")` of
fluid.hpp:1164-1166, an unknown solver name raises ValueError like the
runtime_error of hydro2d.hpp:247.

Every call goes to libhydrogpu.so; there is no CPU fallback.
"""
import ctypes

import numpy as np

from . import capi

FACE_FIELDS = (capi.HG_FIELD_VOLUME_FLUX, capi.HG_FIELD_VOLUME_FLUX_ITER,
               capi.HG_FIELD_PRESSURE_FACE_COEF)


class HydroError(RuntimeError):
    pass


class HydroNaN(HydroError):
    pass


class Unsupported(HydroError):
    pass


def _raise(lib, ctx, rc, what):
    if rc == capi.HG_OK:
        return
    msg = lib.hg_last_error(ctx).decode() if ctx else ""
    text = f"{what}: {capi.STATUS_NAMES.get(rc, rc)} {msg}".strip()
    if rc == capi.HG_ERR_NAN:
        raise HydroNaN(msg or "NaN")
    if rc == capi.HG_ERR_UNSUPPORTED:
        raise Unsupported(text)
    raise HydroError(text)


class Hydro:
    """One GPU-resident hydro2d case."""

    def __init__(self, config):
        self.lib = capi.load()
        if isinstance(config, str):
            config = capi.parse_config(config, lib=self.lib)
        self.cfg = config
        self.ctx = ctypes.c_void_p()
        rc = self.lib.hg_create(ctypes.byref(config), ctypes.byref(self.ctx))
        if rc != capi.HG_OK:
            self.ctx = ctypes.c_void_p()
            _raise(self.lib, None, rc, "hg_create")
        self.nc = self.lib.hg_num_cells(self.ctx)
        self.nf = self.lib.hg_num_faces(self.ctx)

    def close(self):
        if self.ctx:
            self.lib.hg_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- fields
    def get(self, field):
        n = self.nf if field in FACE_FIELDS else self.nc
        out = np.empty(n, dtype=np.float64)
        _raise(self.lib, self.ctx, self.lib.hg_get_field(self.ctx, field, capi.dptr(out), n),
               "hg_get_field")
        return out

    def set(self, field, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        _raise(self.lib, self.ctx, self.lib.hg_set_field(self.ctx, field, capi.dptr(a), a.size),
               "hg_set_field")

    def refresh(self):
        _raise(self.lib, self.ctx, self.lib.hg_refresh(self.ctx), "hg_refresh")

    # ---- stepping
    def step(self):
        st = capi.HgStepStats()
        _raise(self.lib, self.ctx, self.lib.hg_step(self.ctx, ctypes.byref(st)), "hg_step")
        return st

    def start_step(self):
        _raise(self.lib, self.ctx, self.lib.hg_fluid_start_step(self.ctx), "start_step")

    def iterate(self):
        st = capi.HgIterStats()
        _raise(self.lib, self.ctx, self.lib.hg_fluid_iterate(self.ctx, ctypes.byref(st)),
               "iterate")
        return st

    def is_converged(self):
        return bool(self.lib.hg_fluid_is_converged(self.ctx))

    def finish_step(self):
        _raise(self.lib, self.ctx, self.lib.hg_fluid_finish_step(self.ctx), "finish_step")

    def advect_substep(self):
        _raise(self.lib, self.ctx, self.lib.hg_advect_substep(self.ctx), "advect_substep")

    def update_properties(self):
        _raise(self.lib, self.ctx, self.lib.hg_update_properties(self.ctx), "update_properties")

    def stats(self):
        st = capi.HgStats()
        _raise(self.lib, self.ctx, self.lib.hg_calc_stats(self.ctx, ctypes.byref(st)),
               "calc_stats")
        return st

    def timing(self):
        f, a, p = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self.lib.hg_last_step_timing(self.ctx, ctypes.byref(f), ctypes.byref(a), ctypes.byref(p))
        return {"step.fluid": f.value, "step.advection": a.value,
                "step.fluid_properties": p.value}


def stencil_solve(nx, ny, s, w, c, e, n, k, solver, tol=0.0, max_iters=100, omega=1.0,
                  precision=0):
    """LinearSolver::Solve on 5-point rows, on the GPU."""
    lib = capi.load()
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (s, w, c, e, n, k)]
    x = np.zeros(nx * ny)
    opts = capi.HgSolveOpts(solver=solver if isinstance(solver, int) else capi.SOLVERS[solver],
                            tolerance=tol, max_iters=max_iters, relaxation_factor=omega,
                            precision=precision)
    st = capi.HgSolveStats()
    rc = lib.hg_stencil_solve(nx, ny, *[capi.dptr(a) for a in arrs], capi.dptr(x),
                              ctypes.byref(opts), ctypes.byref(st))
    _raise(lib, None, rc, "hg_stencil_solve")
    return x, st


class Poisson:
    """Resident variable-coefficient pressure operator (face form)."""

    def __init__(self, nx, ny, kx, ky):
        self.lib = capi.load()
        self.nx, self.ny = nx, ny
        self._kx = np.ascontiguousarray(kx, dtype=np.float64)
        self._ky = np.ascontiguousarray(ky, dtype=np.float64)
        self.ps = ctypes.c_void_p()
        rc = self.lib.hg_poisson_create(nx, ny, capi.dptr(self._kx), capi.dptr(self._ky),
                                        ctypes.byref(self.ps))
        _raise(self.lib, None, rc, "hg_poisson_create")

    def solve(self, rhs, solver="cg", tol=1e-8, max_iters=100000, precision=0):
        b = np.ascontiguousarray(rhs, dtype=np.float64)
        p = np.zeros(self.nx * self.ny)
        opts = capi.HgSolveOpts(solver=capi.SOLVERS[solver], tolerance=tol,
                                max_iters=max_iters, relaxation_factor=1.0, precision=precision)
        st = capi.HgSolveStats()
        rc = self.lib.hg_poisson_solve(self.ps, capi.dptr(b), capi.dptr(p), ctypes.byref(opts),
                                       ctypes.byref(st))
        if rc not in (capi.HG_OK, capi.HG_ERR_NOT_CONVERGED):
            _raise(self.lib, None, rc, "hg_poisson_solve")
        return p, st

    def close(self):
        if self.ps:
            self.lib.hg_poisson_destroy(self.ps)
            self.ps = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
