"""The 3D step on the GPU (hg3_*, include/hydro_gpu3d.h) against the
reference's own 3D module (tests/golden/ref3d_*, written by
oracle/_ref/hydro_ref3d) and the 3D oracle.

* With the jacobi solvers every field of every dumped step equals the
  reference's bit for bit, with the same SIMPLE iteration counts: cavity
  12x10x8 (one iteration per step and the loop to convergence), the
  two-phase dam 16x4x8, and cavity 32^3 over 60 steps (SHA-256).
* The reference's default solvers (gauss_seidel, lu) are exact-order
  recurrences; the 3D path refuses them at hg3_create.
* cg (FP64 Jacobi-PCG to pressure_tolerance_gpu): cavity 24^3 against the
  oracle's cg, relative L2 per field.
* fcg_mg (FP64 flexible CG with the FP32 V-cycle, no reference
  counterpart): cavity 24^3 and the two-phase dam 32x8x16 (density ratio
  1000) against the oracle's cg, both to a 1e-12 residual, L2 differences
  relative to the velocity field's and the pressure's norms (p up to a
  constant: the all-wall operator is singular), with a
  small fraction of cg's iterations; grids that do not halve to a small
  coarsest level are refused.
"""
import numpy as np
import pytest

from hydro_amd import capi, workloads
from hydro_amd.hydro import Hydro3, Unsupported
from oracle import pyoracle
from tests.test_oracle3d import check_case, check_long

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["ref3d_cavity12x10x8", "ref3d_cavity12x10x8_loop",
                                  "ref3d_dam16x4x8"])
def test_gpu3d_matches_reference_bitwise(gpu, name):
    print(f"case: {name}", flush=True)
    check_case(name, Hydro3).close()


def test_gpu3d_long_cavity32_bitwise(gpu):
    print("case: cavity 32^3, 60 steps", flush=True)
    check_long(Hydro3).close()


def test_gpu3d_refuses_exact_order_solvers(gpu):
    with pytest.raises(Unsupported):
        Hydro3(workloads.cavity3d(10, 8, 6, solver_p="gauss_seidel", solver_v="lu"))
    with pytest.raises(Unsupported):
        Hydro3(workloads.cavity3d(8, 8, 8) + 'set string condition_left "inlet 1 0 0"\n')


def test_gpu3d_cg_against_oracle(gpu):
    text = workloads.cavity3d(24, 24, 24, solver_p="cg", extra=(
        "set double pressure_tolerance_gpu 1e-10\nset int pressure_max_iters_gpu 2000\n"))
    print("case: cavity 24^3 cg, 5 steps", flush=True)
    o = pyoracle.Oracle3(capi.parse_config3(text))
    with Hydro3(text) as h:
        for _ in range(5):
            sg, so = h.step(), o.step()
            assert sg.simple_iterations == so.simple_iterations
        errs = {}
        for k, fid in (("u", capi.HG_FIELD_VELOCITY_X), ("v", capi.HG_FIELD_VELOCITY_Y),
                       ("w", capi.HG_FIELD_VELOCITY_Z), ("p", capi.HG_FIELD_PRESSURE)):
            a, b = h.get(fid), o.get(fid)
            if k == "p":
                a, b = a - a.mean(), b - b.mean()
            errs[k] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    o.close()
    print("GPU cg vs oracle cg: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    for k, v in errs.items():
        assert v <= 1e-8, (k, v)


def _rel_errors(h, o):
    """Per field: the L2 difference relative to the L2 norm of the whole
    velocity field (u, v, w together: a component that stays at round-off
    level, as v across the depth-uniform dam, has no scale of its own) or of
    the pressure (mean-removed)."""
    vel = ("u", capi.HG_FIELD_VELOCITY_X), ("v", capi.HG_FIELD_VELOCITY_Y), \
        ("w", capi.HG_FIELD_VELOCITY_Z)
    ref = {k: o.get(f) for k, f in vel}
    scale = float(np.sqrt(sum(np.sum(b * b) for b in ref.values())))
    errs = {k: float(np.linalg.norm(h.get(f) - ref[k]) / scale) for k, f in vel}
    a, b = h.get(capi.HG_FIELD_PRESSURE), o.get(capi.HG_FIELD_PRESSURE)
    a, b = a - a.mean(), b - b.mean()
    errs["p"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    return errs


@pytest.mark.parametrize("case", ["cavity24", "dam32x8x16"])
def test_gpu3d_fcg_mg_against_oracle(gpu, case):
    tight = "set double pressure_tolerance_gpu 1e-12\nset int pressure_max_iters_gpu 4000\n"
    if case == "cavity24":
        make = lambda sp: workloads.cavity3d(24, 24, 24, solver_p=sp, extra=tight)
        steps = 5
    else:
        make = lambda sp: workloads.dam3d(32, 8, 16, solver_p=sp, extra=tight)
        steps = 4
    print(f"case: {case} fcg_mg against the oracle's cg, {steps} steps", flush=True)
    o = pyoracle.Oracle3(capi.parse_config3(make("cg")))
    it_mg, it_cg = 0, 0
    with Hydro3(make("fcg_mg")) as h:
        for _ in range(steps):
            sg, so = h.step(), o.step()
            assert sg.simple_iterations == so.simple_iterations
            it_mg += sg.pressure_solve_iters_total
            it_cg += so.pressure_solve_iters_total
        errs = _rel_errors(h, o)
    o.close()
    print(f"GPU fcg_mg vs oracle cg ({it_mg} against {it_cg} pressure iterations): " +
          ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    for k, v in errs.items():
        assert v <= 1e-8, (k, v)
    assert 0 < it_mg * 8 < it_cg


def test_gpu3d_fcg_mg_refuses_grids_without_a_small_coarsest_level(gpu):
    # 25^3 is odd at once: its coarsest level would be all 15625 cells
    with pytest.raises(Unsupported, match="fcg_mg"):
        Hydro3(workloads.cavity3d(25, 25, 25, solver_p="fcg_mg"))
    # 26 x 26 x 26 halves once to 13^3 = 2197 cells: accepted
    Hydro3(workloads.cavity3d(26, 26, 26, solver_p="fcg_mg")).close()


@pytest.mark.parametrize("solver", ["cg", "fcg_mg"])
def test_gpu3d_iteration_cap_reports_not_converged(gpu, solver):
    """A pressure solve cut off by pressure_max_iters_gpu clears the step's
    pressure_converged; the same case with the default cap sets it."""
    print(f"case: cavity 16^3 {solver}, pressure_max_iters_gpu 1", flush=True)
    capped = workloads.cavity3d(16, 16, 16, solver_p=solver,
                                extra="set int pressure_max_iters_gpu 1\n")
    with Hydro3(capped) as h:
        st = h.step()
        assert st.pressure_converged == 0
    with Hydro3(workloads.cavity3d(16, 16, 16, solver_p=solver)) as h:
        st = h.step()
        assert st.pressure_converged == 1
