"""hg_stencil_solve_noalloc (include/hydro_noalloc.h) against hg_stencil_solve.

The allocation-free entry point runs the same row solvers on device arrays,
with its work arrays in the library's file-scope scratch array.  For every
solver it must give the same solution bit for bit, and the same iteration
count, as hg_stencil_solve on the same rows, on two calls in a row with the
second on a smaller grid (the scratch array keeps the first call's contents,
so whatever must start at zero has to be cleared by the call itself).  A
request beyond the scratch array's size is refused with HG_ERR_BAD_ARG.
"""
import ctypes

import numpy as np
import pytest

from hydro_amd import capi
from hydro_amd.hydro import stencil_solve

pytestmark = pytest.mark.gpu

SOLVERS = ["lu", "lu_rb", "gauss_seidel", "jacobi", "lu_relaxed"]


class _Hip:
    def __init__(self):
        try:
            self.lib = ctypes.CDLL("libamdhip64.so")
        except OSError:
            self.lib = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int]
        self.ptrs = []

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.ptrs.append(p)
        assert self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, n):
        out = np.empty(n, dtype=np.float64)
        assert self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.ptrs:
            self.lib.hipFree(p)
        self.ptrs = []


def _noalloc(lib):
    return lib.hg_stencil_solve_noalloc  # declared by capi


def _system(nx, ny, seed):
    rng = np.random.default_rng(seed)
    n = nx * ny
    s, w, e, nn = (-rng.uniform(0.5, 1.0, n) for _ in range(4))
    i = np.arange(n) % nx
    j = np.arange(n) // nx
    s[j == 0] = 0
    w[i == 0] = 0
    e[i == nx - 1] = 0
    nn[j == ny - 1] = 0
    c = -(s + w + e + nn) + rng.uniform(0.1, 0.5, n)
    k = rng.normal(size=n)
    return s, w, c, e, nn, k


@pytest.mark.parametrize("solver", SOLVERS)
def test_noalloc_matches_stencil_solve(gpu, solver):
    lib = capi.load()
    fn = _noalloc(lib)
    hip = _Hip()
    sid = capi.SOLVERS[solver]
    omega = {"jacobi": 0.8, "lu_relaxed": 0.5}.get(solver, 1.5)
    try:
        # two calls in a row, the second on a smaller grid
        for (nx, ny, seed) in ((48, 40, 3), (24, 20, 4)):
            print(f"case: noalloc {solver} {nx}x{ny}", flush=True)
            rows = _system(nx, ny, seed)
            x_ref, st_ref = stencil_solve(nx, ny, *rows, sid, tol=1e-6, max_iters=300, omega=omega)
            dev = [hip.upload(a) for a in rows]
            dx = hip.upload(np.full(nx * ny, np.nan))
            opts = capi.HgSolveOpts(solver=sid, tolerance=1e-6, max_iters=300,
                                    relaxation_factor=omega, precision=0)
            st = capi.HgSolveStats()
            rc = fn(nx, ny, *dev, dx, ctypes.byref(opts), ctypes.byref(st), None)
            assert rc == capi.HG_OK, lib.hg_last_error(None)
            x = hip.download(dx, nx * ny)
            assert np.array_equal(x, x_ref), (solver, nx, ny)
            assert st.iterations == st_ref.iterations
            assert st.converged == st_ref.converged
            hip.free()
    finally:
        hip.free()


def test_noalloc_refuses_sizes_beyond_its_scratch(gpu):
    lib = capi.load()
    fn = _noalloc(lib)
    hip = _Hip()
    try:
        nx, ny = 1024, 1024  # 4 arrays of 8 MiB: beyond the 16 MiB scratch array
        one = hip.upload(np.zeros(1))
        opts = capi.HgSolveOpts(solver=capi.SOLVERS["jacobi"], tolerance=1e-6, max_iters=10,
                                relaxation_factor=0.8, precision=0)
        st = capi.HgSolveStats()
        rc = fn(nx, ny, *([one] * 7), ctypes.byref(opts), ctypes.byref(st), None)
        assert rc == capi.HG_ERR_BAD_ARG
        assert b"scratch" in lib.hg_last_error(None)
    finally:
        hip.free()
