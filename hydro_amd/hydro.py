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


def _raise(lib, ctx, rc, what, create=False):
    if rc == capi.HG_OK:
        return
    if ctx:
        msg = lib.hg_last_error(ctx).decode()
    else:  # hg_last_error(NULL): the last failed hg_create's message
        msg = lib.hg_last_error(None).decode() if create else ""
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
            _raise(self.lib, None, rc, "hg_create", create=True)
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

    def progress(self):
        """Fluid time, steps finished, "s", convergence indicator, meshpos."""
        pr = capi.HgProgress()
        _raise(self.lib, self.ctx, self.lib.hg_get_progress(self.ctx, ctypes.byref(pr)),
               "hg_get_progress")
        return pr

    def iteration_history(self):
        """Rs of the SIMPLE iterations of the current (or last) step."""
        n = ctypes.c_int32()
        _raise(self.lib, self.ctx,
               self.lib.hg_get_iteration_history(self.ctx, None, 0, ctypes.byref(n)),
               "hg_get_iteration_history")
        rs = np.zeros(max(n.value, 1))
        _raise(self.lib, self.ctx,
               self.lib.hg_get_iteration_history(self.ctx, capi.dptr(rs), n.value,
                                                 ctypes.byref(n)),
               "hg_get_iteration_history")
        return rs[:n.value]

    def open_output(self, config):
        """An output session on this case (hydro::InitOutput): `config` is
        .hydroconf text or an HgOutputConfig."""
        return Output(self, config)

    def train_pressure_net(self, rhs_count=None, mg_iters=None, steps=None, lr=None,
                           seed=None, flags=None):
        """Trains the learned pressure preconditioner on the operator of the
        last pressure solve (linear_solver_pressure fcg_learned)."""
        opts = train_opts(rhs_count, mg_iters, steps, lr, seed, flags)
        st = capi.HgTrainStats()
        rc = self.lib.hg_train_pressure_net(self.ctx, ctypes.byref(opts), ctypes.byref(st))
        _raise(self.lib, self.ctx, rc, "hg_train_pressure_net")
        return st

    def time_kernel(self, kernel, reps=20):
        """Average device ms of one launch of a step kernel on the current state."""
        ms = ctypes.c_double()
        rc = self.lib.hg_time_kernel(self.ctx, kernel.encode(), reps, ctypes.byref(ms))
        _raise(self.lib, self.ctx, rc, "hg_time_kernel")
        return ms.value

    def timing(self):
        f, a, p = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self.lib.hg_last_step_timing(self.ctx, ctypes.byref(f), ctypes.byref(a), ctypes.byref(p))
        return {"step.fluid": f.value, "step.advection": a.value,
                "step.fluid_properties": p.value}


class Output:
    """The module's output sessions over a Hydro (hg_output_*): VTS/PVD
    frames, the scalar .dat table and the iteration history, written by
    `write_results(force)` as hydro::write_results (hydro2d.hpp:1623-1652)."""

    def __init__(self, hydro, config):
        self.lib = hydro.lib
        self.hydro = hydro
        if isinstance(config, str):
            config = capi.parse_output_config(config, lib=self.lib)
        self.oc = config
        self.h = ctypes.c_void_p()
        rc = self.lib.hg_output_open(hydro.ctx, ctypes.byref(config), ctypes.byref(self.h))
        if rc != capi.HG_OK:
            self.h = ctypes.c_void_p()
            self._fail(rc, "hg_output_open", None)

    def _fail(self, rc, what, h):
        msg = self.lib.hg_output_last_error(h).decode()
        text = f"{what}: {capi.STATUS_NAMES.get(rc, rc)} {msg}".strip()
        if rc == capi.HG_ERR_UNSUPPORTED:
            raise Unsupported(text)
        raise HydroError(text)

    def write_results(self, force=False):
        rc = self.lib.hg_output_write_results(self.h, 1 if force else 0)
        if rc != capi.HG_OK:
            self._fail(rc, "hg_output_write_results", self.h)

    def close(self):
        if self.h:
            h, self.h = self.h, ctypes.c_void_p()
            rc = self.lib.hg_output_close(h)
            if rc != capi.HG_OK:
                self._fail(rc, "hg_output_close", None)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


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

    def time_kernel(self, kernel, reps=20):
        """Average device ms of one launch of a pressure-solve kernel."""
        ms = ctypes.c_double()
        rc = self.lib.hg_poisson_time_kernel(self.ps, kernel.encode(), reps, ctypes.byref(ms))
        _raise(self.lib, None, rc, "hg_poisson_time_kernel")
        return ms.value

    def train(self, rhs_count=None, mg_iters=None, steps=None, lr=None, seed=None, flags=None):
        """Trains the learned preconditioner on this operator; returns the
        hg_train_stats (pairs, steps, loss_first, loss_last, ms)."""
        opts = train_opts(rhs_count, mg_iters, steps, lr, seed, flags)
        st = capi.HgTrainStats()
        rc = self.lib.hg_poisson_train(self.ps, ctypes.byref(opts), ctypes.byref(st))
        _raise(self.lib, None, rc, "hg_poisson_train")
        return st

    def close(self):
        if self.ps:
            self.lib.hg_poisson_destroy(self.ps)
            self.ps = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def train_opts(rhs_count=None, mg_iters=None, steps=None, lr=None, seed=None, flags=None):
    """hg_train_opts: library defaults, overridden where given."""
    lib = capi.load()
    o = capi.HgTrainOpts()
    lib.hg_train_opts_default(ctypes.byref(o))
    for name, v in (("rhs_count", rhs_count), ("mg_iters", mg_iters), ("steps", steps),
                    ("lr", lr), ("seed", seed), ("flags", flags)):
        if v is not None:
            setattr(o, name, v)
    return o


class Net:
    """The learned preconditioner's network alone, on an nx x ny grid."""

    def __init__(self, nx, ny):
        self.lib = capi.load()
        self.nx, self.ny = nx, ny
        self.h = ctypes.c_void_p()
        _raise(self.lib, None, self.lib.hg_net_create(nx, ny, ctypes.byref(self.h)),
               "hg_net_create")
        self.num_weights = int(self.lib.hg_net_num_weights(self.h))

    def _f32(self, a, n):
        a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
        if a.size != n:
            raise ValueError(f"expected {n} values, got {a.size}")
        return a

    def configure(self, linear=False, tied=False):
        flags = (capi.HG_TRAIN_LINEAR if linear else 0) | (capi.HG_TRAIN_TIED if tied else 0)
        _raise(self.lib, None, self.lib.hg_net_configure(self.h, flags), "hg_net_configure")

    def set_weights(self, w):
        w = self._f32(w, self.num_weights)
        _raise(self.lib, None, self.lib.hg_net_set_weights(self.h, capi.fptr(w)),
               "hg_net_set_weights")

    def get_weights(self):
        w = np.zeros(self.num_weights, np.float32)
        _raise(self.lib, None, self.lib.hg_net_get_weights(self.h, capi.fptr(w)),
               "hg_net_get_weights")
        return w

    def forward(self, x):
        n = self.nx * self.ny
        x = self._f32(x, n)
        y = np.zeros(n, np.float32)
        _raise(self.lib, None, self.lib.hg_net_forward(self.h, capi.fptr(x), capi.fptr(y)),
               "hg_net_forward")
        return y.reshape(self.ny, self.nx)

    def gradient(self, x, gy):
        n = self.nx * self.ny
        x, gy = self._f32(x, n), self._f32(gy, n)
        gw = np.zeros(self.num_weights, np.float32)
        _raise(self.lib, None, self.lib.hg_net_gradient(self.h, capi.fptr(x), capi.fptr(gy),
                                                        capi.fptr(gw)), "hg_net_gradient")
        return gw

    def train_step(self, x, t, lr):
        n = self.nx * self.ny
        x, t = self._f32(x, n), self._f32(t, n)
        loss = ctypes.c_double()
        _raise(self.lib, None, self.lib.hg_net_train_step(self.h, capi.fptr(x), capi.fptr(t),
                                                          lr, ctypes.byref(loss)),
               "hg_net_train_step")
        return loss.value

    def time(self, reps=10):
        v = [ctypes.c_double() for _ in range(4)]
        _raise(self.lib, None, self.lib.hg_net_time(self.h, reps, *[ctypes.byref(a) for a in v]),
               "hg_net_time")
        return {"forward_ms": v[0].value, "train_ms": v[1].value,
                "forward_flops": v[2].value, "train_flops": v[3].value}

    def close(self):
        if self.h:
            self.lib.hg_net_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_run(cases, steps, threads=0):
    """Steps every Hydro in `cases` `steps` times concurrently (hg_batch_run);
    returns (last step stats per case, wall ms)."""
    lib = capi.load()
    n = len(cases)
    arr = (ctypes.c_void_p * n)(*[h.ctx for h in cases])
    stats = (capi.HgStepStats * n)()
    ms = ctypes.c_double()
    rc = lib.hg_batch_run(ctypes.cast(arr, ctypes.POINTER(ctypes.c_void_p)), n, steps, threads,
                          stats, ctypes.byref(ms))
    _raise(lib, None, rc, "hg_batch_run")
    return list(stats), ms.value


def _raise3(lib, ctx, rc, what):
    if rc == capi.HG_OK:
        return
    msg = lib.hg3_last_error(ctx if ctx else None).decode()
    text = f"{what}: {capi.STATUS_NAMES.get(rc, rc)} {msg}".strip()
    if rc == capi.HG_ERR_NAN:
        raise HydroNaN(msg or "NaN")
    if rc == capi.HG_ERR_UNSUPPORTED:
        raise Unsupported(text)
    raise HydroError(text)


class Hydro3:
    """One GPU-resident hydro3d case (include/hydro_gpu3d.h): the same
    step as Hydro on the structured 3D mesh."""

    def __init__(self, config):
        self.lib = capi.load()
        if isinstance(config, str):
            config = capi.parse_config3(config, lib=self.lib)
        self.cfg = config
        self.ctx = ctypes.c_void_p()
        rc = self.lib.hg3_create(ctypes.byref(config), ctypes.byref(self.ctx))
        if rc != capi.HG_OK:
            self.ctx = ctypes.c_void_p()
            _raise3(self.lib, None, rc, "hg3_create")
        nx, ny, nz = config.base.nx, config.base.ny, config.nz
        self.nc = nx * ny * nz
        self.nf = (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1)

    def close(self):
        if self.ctx:
            self.lib.hg3_destroy(self.ctx)
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

    def get(self, field):
        n = self.nf if field == capi.HG_FIELD_VOLUME_FLUX else self.nc
        out = np.empty(n, dtype=np.float64)
        _raise3(self.lib, self.ctx, self.lib.hg3_get_field(self.ctx, field, capi.dptr(out), n),
                "hg3_get_field")
        return out

    def set(self, field, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        _raise3(self.lib, self.ctx, self.lib.hg3_set_field(self.ctx, field, capi.dptr(v), v.size),
                "hg3_set_field")

    def step(self):
        st = capi.HgStepStats()
        _raise3(self.lib, self.ctx, self.lib.hg3_step(self.ctx, ctypes.byref(st)), "hg3_step")
        return st

