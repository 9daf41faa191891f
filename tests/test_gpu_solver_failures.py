"""The pressure solvers' failure paths.

* FCG's fallback: when the preconditioner breaks the descent property
  ((r, z) <= 0 or (p, A p) <= 0), flexible CG continues with FP64 Jacobi-PCG
  from its current iterate to the same tolerance.  The check knob
  fcg_negate_precond_gpu makes FCG apply minus its V-cycle, which fails the
  check on the first iteration: the solve must still reach the 1e-8 true
  residual, with about Jacobi-PCG's iteration count, and give the pressure
  of the solve without the knob (and of the CPU oracle's CG) up to the
  solver tolerance.
* The iteration cap: with pressure_max_iters_gpu 1 a solve stops unconverged,
  and the iteration statistics say so (pressure_converged 0, the residual
  above the tolerance), as the reference's solvers only print their final
  difference when they hit their limit (linear.hpp:710-712).
"""
import numpy as np
import pytest

from hydro_amd import capi, workloads
from hydro_amd.hydro import Hydro
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def rel_mean_removed(a, b):
    a = a - a.mean()
    b = b - b.mean()
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def first_iteration(h):
    h.start_step()
    st = h.iterate()
    h.finish_step()
    return st, h.get(capi.HG_FIELD_PRESSURE)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fcg_fallback_reaches_the_tolerance(gpu, precision):
    text = workloads.dam(64, 32, solver_p="fcg_mg", solver_v="lu_rb",
                         extra=f"set string precision_gpu {precision}\n")
    print(f"case: dam 64x32, fcg_mg {precision}, with and without the negated V-cycle",
          flush=True)
    with Hydro(text) as a, Hydro(text + "set bool fcg_negate_precond_gpu 1\n") as b:
        sa, pa = first_iteration(a)
        sb, pb = first_iteration(b)
    print(f"  plain: {sa.pressure_solve_iters} iterations, residual {sa.pressure_solve_diff:.2e}; "
          f"fallback: {sb.pressure_solve_iters} iterations, residual "
          f"{sb.pressure_solve_diff:.2e}", flush=True)
    assert sa.pressure_converged == 1 and sb.pressure_converged == 1
    assert sa.pressure_solve_diff <= 1e-8 and sb.pressure_solve_diff <= 1e-8
    # the fallback is Jacobi-PCG: many more iterations than the V-cycle's
    assert sb.pressure_solve_iters > 4 * sa.pressure_solve_iters
    assert rel_mean_removed(pb, pa) <= 1e-6
    if precision == "fp64":
        o = pyoracle.Oracle(capi.parse_config(workloads.dam(64, 32, solver_p="cg",
                                                             solver_v="lu_rb")))
        o.start_step()
        o.iterate()
        o.finish_step()
        assert rel_mean_removed(pb, o.get(capi.HG_FIELD_PRESSURE)) <= 1e-6


@pytest.mark.parametrize("solver,precision", [("cg", "fp64"), ("cg", "mixed"),
                                              ("fcg_mg", "mixed")])
def test_iteration_cap_reports_not_converged(gpu, solver, precision):
    text = workloads.dam(64, 32, solver_p=solver, solver_v="lu_rb",
                         extra=f"set string precision_gpu {precision}\n"
                               "set int pressure_max_iters_gpu 1\n")
    print(f"case: dam 64x32, {solver} {precision}, at most 1 pressure iteration", flush=True)
    with Hydro(text) as h:
        h.start_step()
        st = h.iterate()
    print(f"  {st.pressure_solve_iters} iterations, residual {st.pressure_solve_diff:.2e}, "
          f"converged {st.pressure_converged}", flush=True)
    assert st.pressure_converged == 0
    assert st.pressure_solve_diff > 1e-8
    assert 1 <= st.pressure_solve_iters <= 2
