"""The C-ABI library and its host logic, without a GPU.

* libhydrogpu.so loads and exports every function include/hydro_gpu.h,
  include/hydro_output.h, include/hydro_noalloc.h and include/hydro_gpu3d.h
  declare (no compute calls here).
* The .hydroconf reader keeps the reference's defaults and semantics.
* Calls that need a device fail loudly when none is visible.
"""
import ctypes
import os
import re

import pytest

from hydro_amd import capi
from tests.conftest import gpu_available

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h)
           for h in ("hydro_gpu.h", "hydro_output.h", "hydro_noalloc.h", "hydro_gpu3d.h")]


def declared_functions():
    text = "".join(open(h).read() for h in HEADERS)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hg3?_[a-z0-9_]+)\s*\(", text)))


def test_header_and_python_mirror_agree():
    assert declared_functions() == sorted(capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_defaults_are_general_hydroconf():
    cfg = capi.default_config()
    assert (cfg.nx, cfg.ny) == (100, 100)
    assert cfg.dt == 0.01
    assert cfg.linear_solver_pressure == capi.SOLVERS["gauss_seidel"]
    assert cfg.linear_solver_velocity == capi.SOLVERS["lu"]
    assert (cfg.lu_relaxed_tolerance, cfg.lu_relaxed_num_iters_limit,
            cfg.lu_relaxed_relaxation_factor) == (1e-3, 1000, 1.9)
    assert (cfg.convergence_tolerance, cfg.num_iterations_limit) == (1e-2, 10)
    assert (cfg.velocity_relaxation_factor, cfg.pressure_relaxation_factor) == (0.8, 0.9)
    assert cfg.advection_dt_factor == 0.1
    assert cfg.density_smooth_times == 2 and cfg.initial_volume_fraction_smooth_times == 2
    assert cfg.unsupported_flags == 0


def test_parse_reference_sample_parameter_dump():
    text = open(os.path.join(ROOT, "tests", "golden", "cavity_sample.hydroconf")).read()
    cfg = capi.parse_config(text)
    assert (cfg.nx, cfg.ny) == (64, 64)
    assert cfg.dt == 100000.0
    assert cfg.viscosity[0] == 0.0003125
    assert tuple(cfg.bc_velocity[capi_side("top")]) == (1.0, 0.0)
    assert cfg.lu_relaxed_num_iters_limit == 100 and cfg.lu_relaxed_tolerance == 0.0
    assert cfg.num_iterations_limit == 1000
    assert cfg.unsupported_flags == 0


def capi_unsup(name):
    return {"radiation": 1 << 3}[name]


def capi_side(name):
    return {"left": 0, "right": 1, "bottom": 2, "top": 3}[name]


def test_parse_overrides_and_del():
    cfg = capi.parse_config("""
set int Nx 7
set vect pressure_fixed_point (0.5, 0.25, 0)
set double pressure_fixed_value 3
set string linear_solver_pressure cg
set string precision_gpu mixed
set string condition_left "wall 0 2 0"
""")
    assert cfg.nx == 7 and cfg.has_pressure_fixed_point
    assert tuple(cfg.pressure_fixed_point) == (0.5, 0.25)
    assert cfg.linear_solver_pressure == capi.SOLVERS["cg"]
    assert cfg.precision_gpu == capi.HG_PREC_MIXED
    assert tuple(cfg.bc_velocity[0]) == (0.0, 2.0)
    cfg = capi.parse_config("del pressure_fixed_point\n", cfg)
    assert not cfg.has_pressure_fixed_point


def test_parse_errors_like_the_reference():
    with pytest.raises(ValueError, match="Unknown linear solver"):
        capi.parse_config("set string linear_solver_pressure pardiso\n")
    with pytest.raises(ValueError, match="Unknown type"):
        capi.parse_config("set float x 1\n")
    with pytest.raises(ValueError):
        capi.parse_config("set vect A 1\n")


def test_unsupported_features_are_flagged():
    cfg = capi.parse_config("set bool compressible_enable 1\nset bool radiation_enable 1\n")
    assert cfg.unsupported_flags != 0
    cfg = capi.parse_config("set bool compressible_enable 0\n", cfg)
    assert cfg.unsupported_flags == capi_unsup("radiation")
    # variants that run: dt_auto, heat, inlets / outlets, SIMPLER
    cfg = capi.parse_config("set bool dt_auto 1\nset double cfl 0.25\nset bool heat_enable 1\n"
                            "set string condition_left \"inlet 1 0\"\n"
                            "set string condition_right outlet\nset bool simpler 1\n")
    assert cfg.unsupported_flags == 0 and cfg.dt_auto == 1 and cfg.cfl == 0.25
    assert cfg.heat_enable == 1 and cfg.simpler == 1
    assert tuple(cfg.bc_kind) == (1, 2, 0, 0) and tuple(cfg.bc_velocity[0]) == (1.0, 0.0)
    assert capi.default_config().cfl_advection == 0.5


@pytest.mark.skipif(gpu_available(), reason="checks the no-device path")
def test_create_fails_loudly_without_device():
    lib = capi.load()
    cfg = capi.parse_config("set int Nx 8\nset int Ny 8\n")
    ctx = ctypes.c_void_p()
    rc = lib.hg_create(ctypes.byref(cfg), ctypes.byref(ctx))
    assert rc == capi.HG_ERR_HIP and not ctx
