"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product path never uses it.
"""
import ctypes
import os

import numpy as np

from hydro_amd import capi

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")
REF_BIN = os.path.join(ORACLE_DIR, "_ref", "hydro_ref")
REF_BIN_FAST = os.path.join(ORACLE_DIR, "_ref", "hydro_ref_fast")
REF_BIN_3D = os.path.join(ORACLE_DIR, "_ref", "hydro_ref3d")

_P = ctypes.POINTER
_dp = _P(ctypes.c_double)
_LIB = None


def load():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    sig = {
        "or_create": (ctypes.c_int, [_P(capi.HgConfig), _P(vp), ctypes.c_char_p, ctypes.c_size_t]),
        "or_destroy": (None, [vp]),
        "or_num_cells": (ctypes.c_int, [vp]),
        "or_num_faces": (ctypes.c_int, [vp]),
        "or_set_field": (ctypes.c_int, [vp, ctypes.c_int, _dp, ctypes.c_size_t]),
        "or_get_field": (ctypes.c_int, [vp, ctypes.c_int, _dp, ctypes.c_size_t]),
        "or_refresh": (ctypes.c_int, [vp]),
        "or_fluid_start_step": (ctypes.c_int, [vp]),
        "or_fluid_iterate": (ctypes.c_int, [vp, _P(capi.HgIterStats)]),
        "or_fluid_is_converged": (ctypes.c_int, [vp]),
        "or_fluid_finish_step": (ctypes.c_int, [vp]),
        "or_advect_substep": (ctypes.c_int, [vp]),
        "or_update_properties": (ctypes.c_int, [vp]),
        "or_calc_stats": (ctypes.c_int, [vp, _P(capi.HgStats)]),
        "or_step": (ctypes.c_int, [vp, _P(capi.HgStepStats)]),
        "or_timer": (ctypes.c_double, [vp, ctypes.c_char_p]),
        "or_get_system": (ctypes.c_int, [vp, ctypes.c_int] + [_dp] * 7),
        "or_stencil_solve": (ctypes.c_int, [ctypes.c_int, ctypes.c_int] + [_dp] * 7 +
                             [_P(capi.HgSolveOpts), _P(capi.HgSolveStats)]),
        "or_poisson_solve": (ctypes.c_int, [ctypes.c_int, ctypes.c_int] + [_dp] * 4 +
                             [_P(capi.HgSolveOpts), _P(capi.HgSolveStats)]),
        "or_geometry": (ctypes.c_int, [vp] + [_dp] * 8),
        "or_poisson_baseline": (ctypes.c_int, [ctypes.c_int, ctypes.c_int] + [_dp] * 3 +
                                [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                 ctypes.c_int, _dp, _dp, _P(ctypes.c_int), _dp]),
        "or_bdf_coeffs": (None, [ctypes.c_double, ctypes.c_int, _dp]),
        "or3_create": (ctypes.c_int, [_P(capi.Hg3Config), _P(vp), ctypes.c_char_p,
                                      ctypes.c_size_t]),
        "or3_destroy": (None, [vp]),
        "or3_num_cells": (ctypes.c_int, [vp]),
        "or3_num_faces": (ctypes.c_int, [vp]),
        "or3_get_field": (ctypes.c_int, [vp, ctypes.c_int, _dp, ctypes.c_size_t]),
        "or3_set_field": (ctypes.c_int, [vp, ctypes.c_int, _dp, ctypes.c_size_t]),
        "or3_step": (ctypes.c_int, [vp, _P(capi.HgStepStats)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


class Oracle:
    """One oracle case: the CPU counterpart of a libhydrogpu context."""

    def __init__(self, cfg):
        self.lib = load()
        self.ctx = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        rc = self.lib.or_create(ctypes.byref(cfg), ctypes.byref(self.ctx), err, 512)
        if rc != 0:
            raise RuntimeError(f"or_create: {err.value.decode()} ({rc})")
        self.nc = self.lib.or_num_cells(self.ctx)
        self.nf = self.lib.or_num_faces(self.ctx)
        self.cfg = cfg

    def close(self):
        if self.ctx:
            self.lib.or_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} -> {capi.STATUS_NAMES.get(rc, rc)}")

    def size_of(self, field):
        return self.nf if field in (capi.HG_FIELD_VOLUME_FLUX, capi.HG_FIELD_VOLUME_FLUX_ITER) else self.nc

    def get(self, field):
        out = np.empty(self.size_of(field), dtype=np.float64)
        self._check(self.lib.or_get_field(self.ctx, field, capi.dptr(out), out.size), "or_get_field")
        return out

    def set(self, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self._check(self.lib.or_set_field(self.ctx, field, capi.dptr(a), a.size), "or_set_field")

    def refresh(self):
        self._check(self.lib.or_refresh(self.ctx), "or_refresh")

    def step(self):
        st = capi.HgStepStats()
        self._check(self.lib.or_step(self.ctx, ctypes.byref(st)), "or_step")
        return st

    def start_step(self):
        self._check(self.lib.or_fluid_start_step(self.ctx), "start")

    def iterate(self):
        st = capi.HgIterStats()
        self._check(self.lib.or_fluid_iterate(self.ctx, ctypes.byref(st)), "iterate")
        return st

    def is_converged(self):
        return bool(self.lib.or_fluid_is_converged(self.ctx))

    def finish_step(self):
        self._check(self.lib.or_fluid_finish_step(self.ctx), "finish")

    def advect_substep(self):
        self._check(self.lib.or_advect_substep(self.ctx), "advect")

    def update_properties(self):
        self._check(self.lib.or_update_properties(self.ctx), "props")

    def stats(self):
        st = capi.HgStats()
        self._check(self.lib.or_calc_stats(self.ctx, ctypes.byref(st)), "stats")
        return st

    def timer(self, region):
        return self.lib.or_timer(self.ctx, region.encode())

    def system(self, which):
        n = self.nc
        arrs = [np.zeros(n) for _ in range(7)]
        self._check(self.lib.or_get_system(self.ctx, which, *[capi.dptr(a) for a in arrs]),
                    "or_get_system")
        return arrs

    def geometry(self):
        cells = [np.zeros(self.nc) for _ in range(3)]
        faces = [np.zeros(self.nf) for _ in range(5)]
        self._check(self.lib.or_geometry(self.ctx, *[capi.dptr(a) for a in cells + faces]),
                    "or_geometry")
        return cells, faces


def stencil_solve(nx, ny, s, w, c, e, n, k, solver, tol=0., max_iters=100, omega=1.,
                  precision=0):
    lib = load()
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (s, w, c, e, n, k)]
    x = np.zeros(nx * ny)
    opts = capi.HgSolveOpts(solver=solver, tolerance=tol, max_iters=max_iters,
                            relaxation_factor=omega, precision=precision)
    st = capi.HgSolveStats()
    rc = lib.or_stencil_solve(nx, ny, *[capi.dptr(a) for a in arrs], capi.dptr(x),
                              ctypes.byref(opts), ctypes.byref(st))
    if rc != 0:
        raise RuntimeError(f"or_stencil_solve -> {rc}")
    return x, st


def poisson_solve(nx, ny, kx, ky, rhs, tol=1e-8, max_iters=100000):
    lib = load()
    kx = np.ascontiguousarray(kx, dtype=np.float64)
    ky = np.ascontiguousarray(ky, dtype=np.float64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    p = np.zeros(nx * ny)
    opts = capi.HgSolveOpts(solver=capi.SOLVERS["cg"], tolerance=tol, max_iters=max_iters,
                            relaxation_factor=1.0, precision=0)
    st = capi.HgSolveStats()
    rc = lib.or_poisson_solve(nx, ny, capi.dptr(kx), capi.dptr(ky), capi.dptr(rhs),
                              capi.dptr(p), ctypes.byref(opts), ctypes.byref(st))
    if rc != 0:
        raise RuntimeError(f"or_poisson_solve -> {rc}")
    return p, st


def poisson_baseline(nx, ny, kx, ky, rhs, method, tol=1e-8, max_iters=100000, omega=1.9,
                     check_every=10):
    """CPU pressure solve of the face-form system to a relative residual:
    method "sor" (the reference's lexicographic SOR, one thread, residual
    checked every `check_every` sweeps) or "pcg" (OpenMP Jacobi-PCG).
    Returns (p, seconds, iterations, relative residual)."""
    lib = load()
    kx = np.ascontiguousarray(kx, dtype=np.float64)
    ky = np.ascontiguousarray(ky, dtype=np.float64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    p = np.zeros(nx * ny)
    sec, rel = ctypes.c_double(), ctypes.c_double()
    it = ctypes.c_int()
    m = {"sor": 0, "pcg": 1}[method]
    rc = lib.or_poisson_baseline(nx, ny, capi.dptr(kx), capi.dptr(ky), capi.dptr(rhs), m, omega,
                                 tol, max_iters, check_every, capi.dptr(p), ctypes.byref(sec),
                                 ctypes.byref(it), ctypes.byref(rel))
    if rc != 0:
        raise RuntimeError(f"or_poisson_baseline -> {rc}")
    return p, sec.value, it.value, rel.value


def bdf_coeffs(dt, second_order=True):
    lib = load()
    out = np.zeros(3)
    lib.or_bdf_coeffs(dt, 1 if second_order else 0, capi.dptr(out))
    return out


class Oracle3:
    """One 3D oracle case (the CPU counterpart of an hg3 context)."""

    def __init__(self, cfg3):
        self.lib = load()
        self.ctx = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        rc = self.lib.or3_create(ctypes.byref(cfg3), ctypes.byref(self.ctx), err, 512)
        if rc != 0:
            raise ValueError(err.value.decode())
        self.nc = self.lib.or3_num_cells(self.ctx)
        self.nf = self.lib.or3_num_faces(self.ctx)

    def close(self):
        if self.ctx:
            self.lib.or3_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def step(self):
        st = capi.HgStepStats()
        rc = self.lib.or3_step(self.ctx, ctypes.byref(st))
        if rc != 0:
            raise RuntimeError(f"or3_step: {rc}")
        return st

    def get(self, field):
        n = self.nf if field == capi.HG_FIELD_VOLUME_FLUX else self.nc
        out = np.empty(n)
        rc = self.lib.or3_get_field(self.ctx, field, out.ctypes.data_as(_dp), n)
        if rc != 0:
            raise ValueError(f"or3_get_field {field}: {rc}")
        return out

    def set(self, field, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        rc = self.lib.or3_set_field(self.ctx, field, v.ctypes.data_as(_dp), v.size)
        if rc != 0:
            raise ValueError(f"or3_set_field {field}: {rc}")


def read_ref_dump3(path):
    """Reads a dump of oracle/_ref/hydro_ref3d (ref_shim.cpp built with
    HR_DIM=3): header nx, ny, nz, num_phases, step; u, v, w, p, flux, then
    alpha_i, pd_i per phase, then T."""
    raw = open(path, "rb").read()
    nx, ny, nz, np_, step = (int(v) for v in np.frombuffer(raw[:20], dtype=np.int32))
    nc = nx * ny * nz
    nf = (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1)
    data = np.frombuffer(raw[20:], dtype=np.float64)
    out = {"nx": nx, "ny": ny, "nz": nz, "num_phases": np_, "step": step}
    off = 0
    for name, n in (("u", nc), ("v", nc), ("w", nc), ("p", nc), ("flux", nf)):
        out[name] = data[off:off + n].copy()
        off += n
    for i in range(np_):
        out[f"alpha{i}"] = data[off:off + nc].copy()
        off += nc
        out[f"pd{i}"] = data[off:off + nc].copy()
        off += nc
    out["T"] = data[off:off + nc].copy()
    return out


def read_ref_dump(path):
    """Reads a dump of oracle/_ref/hydro_ref (see oracle/ref_shim.cpp)."""
    raw = open(path, "rb").read()
    nx, ny, np_, step = np.frombuffer(raw[:16], dtype=np.int32)
    nc = int(nx) * int(ny)
    nf = (int(nx) + 1) * int(ny) + int(nx) * (int(ny) + 1)
    data = np.frombuffer(raw[16:], dtype=np.float64)
    out = {"nx": int(nx), "ny": int(ny), "num_phases": int(np_), "step": int(step)}
    off = 0
    for name, n in (("u", nc), ("v", nc), ("p", nc), ("flux", nf)):
        out[name] = data[off:off + n].copy()
        off += n
    for i in range(int(np_)):
        out[f"alpha{i}"] = data[off:off + nc].copy()
        off += nc
        out[f"pd{i}"] = data[off:off + nc].copy()
        off += nc
    if data.size >= off + nc:  # dumps written since the temperature was added
        out["T"] = data[off:off + nc].copy()
    return out
