"""Pressure Poisson solves on the GPU (variable coefficients, density ratio
1000), judged by the true FP64 residual computed here in numpy."""
import numpy as np
import pytest

from hydro_amd import workloads
from hydro_amd.hydro import Poisson
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def apply(kx, ky, p, nx, ny):
    kx = kx.reshape(ny, nx + 1)
    ky = ky.reshape(ny + 1, nx)
    P = p.reshape(ny, nx)
    d = kx[:, :-1] + kx[:, 1:] + ky[:-1, :] + ky[1:, :]
    y = d * P
    y[:, 1:] -= kx[:, 1:-1] * P[:, :-1]
    y[:, :-1] -= kx[:, 1:-1] * P[:, 1:]
    y[1:, :] -= ky[1:-1, :] * P[:-1, :]
    y[:-1, :] -= ky[1:-1, :] * P[1:, :]
    return y.reshape(-1)


@pytest.fixture(scope="module")
def problem():
    n = 128
    kx, ky, rhs = workloads.poisson_problem(n, seed=3)
    return n, kx, ky, rhs


@pytest.mark.parametrize("solver,precision", [("cg", 0), ("cg", 1), ("fcg_learned", 1),
                                              ("fcg_mg", 1), ("fcg_mg", 0)])
def test_poisson_true_residual(gpu, problem, solver, precision):
    n, kx, ky, rhs = problem
    print(f"case: poisson {n}x{n} {solver} precision {precision}", flush=True)
    ps = Poisson(n, n, kx, ky)
    p, st = ps.solve(rhs, solver=solver, tol=1e-8, precision=precision)
    b = rhs - rhs.mean()
    res = np.linalg.norm(b - apply(kx, ky, p, n, n)) / np.linalg.norm(b)
    assert st.converged and res <= 1.05e-8, (res, st.rel_residual, st.iterations)
    assert abs(res - st.rel_residual) <= 1e-3 * res + 1e-14


def test_poisson_matches_oracle_cg(gpu, problem):
    n, kx, ky, rhs = problem
    ps = Poisson(n, n, kx, ky)
    pg, sg = ps.solve(rhs, solver="cg", tol=1e-12, precision=0)
    po, so = pyoracle.poisson_solve(n, n, kx, ky, rhs, tol=1e-12)
    # both are solutions of the singular Neumann system: compare mean-free parts
    pg = pg - pg.mean()
    po = po - po.mean()
    assert np.linalg.norm(pg - po) <= 1e-8 * np.linalg.norm(po)


def test_multigrid_cuts_iterations(gpu, problem):
    n, kx, ky, rhs = problem
    ps = Poisson(n, n, kx, ky)
    _, s_cg = ps.solve(rhs, solver="cg", tol=1e-8, precision=0)
    _, s_mg = ps.solve(rhs, solver="fcg_learned", tol=1e-8, precision=1)
    print(f"jacobi-pcg {s_cg.iterations} it {s_cg.ms:.2f} ms; fcg+mg {s_mg.iterations} it "
          f"{s_mg.ms:.2f} ms", flush=True)
    assert s_mg.iterations * 10 < s_cg.iterations


def test_fp64_vcycle_same_iterations(gpu, problem):
    """fcg_mg with the V-cycle on FP64 levels (precision 0) against the FP32
    one (precision 1): the same kernels, so FP32 costs the preconditioner no
    strength (iteration counts within one), only halves its bytes."""
    n, kx, ky, rhs = problem
    ps = Poisson(n, n, kx, ky)
    _, s32 = ps.solve(rhs, solver="fcg_mg", tol=1e-8, precision=1)
    _, s64 = ps.solve(rhs, solver="fcg_mg", tol=1e-8, precision=0)
    print(f"fcg_mg FP32 V-cycle {s32.iterations} it {s32.ms:.3f} ms; FP64 V-cycle "
          f"{s64.iterations} it {s64.ms:.3f} ms", flush=True)
    assert s32.converged and s64.converged
    assert abs(s32.iterations - s64.iterations) <= 1
