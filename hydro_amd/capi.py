"""ctypes mirror of include/hydro_gpu.h, include/hydro_output.h and
include/hydro_noalloc.h (the C-ABI
of libhydrogpu.so).

The structures below follow the header field by field; `HG_ABI_VERSION`
guards against a stale library.  Everything here is plumbing: the compute
lives in the HIP kernels behind the C-ABI.
"""
import ctypes
import os

HG_ABI_VERSION = 4
HG_MAX_PHASES = 3

HG_OK, HG_ERR_NAN, HG_ERR_NOT_CONVERGED, HG_ERR_BAD_ARG, HG_ERR_HIP, \
    HG_ERR_UNSUPPORTED, HG_ERR_PARSE, HG_ERR_IO = range(8)
STATUS_NAMES = {0: "HG_OK", 1: "HG_ERR_NAN", 2: "HG_ERR_NOT_CONVERGED",
                3: "HG_ERR_BAD_ARG", 4: "HG_ERR_HIP", 5: "HG_ERR_UNSUPPORTED",
                6: "HG_ERR_PARSE", 7: "HG_ERR_IO"}

SOLVERS = {"lu": 0, "lu_relaxed": 1, "gauss_seidel": 2, "jacobi": 3,
           "lu_rb": 4, "cg": 5, "fcg_learned": 6, "fcg_mg": 7}
HG_PREC_FP64, HG_PREC_MIXED = 0, 1
HG_TRAIN_LINEAR, HG_TRAIN_STANDALONE, HG_TRAIN_TIED = 1, 2, 4

HG_FIELD_VELOCITY_X = 0
HG_FIELD_VELOCITY_Y = 1
HG_FIELD_PRESSURE = 2
HG_FIELD_VOLUME_FLUX = 3
HG_FIELD_PARTIAL_DENSITY_0 = 4
HG_FIELD_VOLUME_FRACTION_0 = 7
HG_FIELD_DENSITY = 10
HG_FIELD_DENSITY_SMOOTH = 11
HG_FIELD_VISCOSITY_SMOOTH = 12
HG_FIELD_FORCE_X = 13
HG_FIELD_FORCE_Y = 14
HG_FIELD_VOLUME_FLUX_ITER = 15
HG_FIELD_PRESSURE_FACE_COEF = 16
HG_FIELD_PRESSURE_RHS = 17
HG_FIELD_TEMPERATURE = 18
# 3D context (hydro_gpu3d.h)
HG_FIELD_VELOCITY_Z = 32
HG_FIELD_FORCE_Z = 33
HG_SIDE_CLOSE, HG_SIDE_FAR = 4, 5

_d2 = ctypes.c_double * 2
_d3 = ctypes.c_double * HG_MAX_PHASES
_i4 = ctypes.c_int32 * 4


class HgConfig(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("a", _d2), ("b", _d2),
        ("dt", ctypes.c_double),
        ("num_phases", ctypes.c_int32),
        ("density", _d3), ("viscosity", _d3),
        ("gravity", _d2), ("force", _d2),
        ("sigma", ctypes.c_double),
        ("bc_kind", _i4),
        ("bc_velocity", _d2 * 4),
        ("initial_velocity", _d2),
        ("has_initial_sin", ctypes.c_int32),
        ("initial_sin_n", _d2),
        ("initial_sin_lambda", ctypes.c_double),
        ("initial_sin_phase", ctypes.c_double),
        ("initial_volume_fraction", _d3),
        ("a1", _d2), ("b1", _d2), ("a2", _d2), ("b2", _d2),
        ("ic", _d2), ("ir", ctypes.c_double),
        ("ic2", _d2), ("ir2", ctypes.c_double),
        ("initial_volume_fraction_smooth_times", ctypes.c_int32),
        ("density_smooth_times", ctypes.c_int32),
        ("viscosity_smooth_times", ctypes.c_int32),
        ("force_smooth_times", ctypes.c_int32),
        ("fluid_enable", ctypes.c_int32),
        ("advection_enable", ctypes.c_int32),
        ("dt_auto", ctypes.c_int32), ("cfl", ctypes.c_double),
        ("cfl_advection", ctypes.c_double),
        ("velocity_relaxation_factor", ctypes.c_double),
        ("pressure_relaxation_factor", ctypes.c_double),
        ("rhie_chow_factor", ctypes.c_double),
        ("convergence_tolerance", ctypes.c_double),
        ("num_iterations_limit", ctypes.c_int32),
        ("time_second_order", ctypes.c_int32),
        ("simpler", ctypes.c_int32),
        ("force_geometric_average", ctypes.c_int32),
        ("guess_extrapolation", ctypes.c_double),
        ("meshvel", _d2),
        ("has_pressure_fixed_point", ctypes.c_int32),
        ("pressure_fixed_point", _d2),
        ("pressure_fixed_value", ctypes.c_double),
        ("linear_solver_velocity", ctypes.c_int32),
        ("linear_solver_pressure", ctypes.c_int32),
        ("lu_relaxed_tolerance", ctypes.c_double),
        ("lu_relaxed_num_iters_limit", ctypes.c_int32),
        ("lu_relaxed_relaxation_factor", ctypes.c_double),
        ("advection_dt_factor", ctypes.c_double),
        ("tvd_split", ctypes.c_int32),
        ("sharp", ctypes.c_double),
        ("heat_enable", ctypes.c_int32), ("linear_solver_heat", ctypes.c_int32),
        ("heat_box_lb", _d2), ("heat_box_rt", _d2), ("conductivity", _d3),
        ("temperature_initial", ctypes.c_double), ("heat_box_temperature", ctypes.c_double),
        ("heat_relaxation_factor", ctypes.c_double),
        ("time_second_order_heat", ctypes.c_int32),
        ("pressure_tolerance_gpu", ctypes.c_double),
        ("pressure_max_iters_gpu", ctypes.c_int32),
        ("precision_gpu", ctypes.c_int32),
        ("learned_levels_gpu", ctypes.c_int32),
        ("learned_channels_gpu", ctypes.c_int32),
        ("pressure_warm_start_gpu", ctypes.c_int32),
        ("fcg_negate_precond_gpu", ctypes.c_int32),
        ("meshvel_output", ctypes.c_int32),
        ("unsupported_flags", ctypes.c_int32),
    ]


class HgIterStats(ctypes.Structure):
    _fields_ = [("iteration", ctypes.c_int32),
                ("convergence_indicator", ctypes.c_double),
                ("velocity_solve_iters", ctypes.c_int32 * 2),
                ("pressure_solve_iters", ctypes.c_int32),
                ("pressure_solve_diff", ctypes.c_double),
                ("pressure_converged", ctypes.c_int32)]


class Hg3Config(ctypes.Structure):
    """hg3_config (include/hydro_gpu3d.h)."""
    _fields_ = [("base", HgConfig), ("nz", ctypes.c_int32),
                ("a_z", ctypes.c_double), ("b_z", ctypes.c_double),
                ("a1_z", ctypes.c_double), ("b1_z", ctypes.c_double),
                ("gravity_z", ctypes.c_double), ("force_z", ctypes.c_double),
                ("initial_velocity_z", ctypes.c_double), ("meshvel_z", ctypes.c_double),
                ("bc_kind", ctypes.c_int32 * 6), ("bc_velocity", (ctypes.c_double * 3) * 6)]


class HgStepStats(ctypes.Structure):
    _fields_ = [("simple_iterations", ctypes.c_int32),
                ("convergence_indicator", ctypes.c_double),
                ("advection_substeps", ctypes.c_int32),
                ("pressure_solve_iters_total", ctypes.c_int64),
                ("time", ctypes.c_double), ("dt", ctypes.c_double),
                ("pressure_solve_diff_max", ctypes.c_double),
                ("pressure_converged", ctypes.c_int32)]


class HgStats(ctypes.Structure):
    _fields_ = [(n, _d3) for n in ("volume", "mass", "pd_min", "pd_max",
                                   "cx", "cy", "vx", "vy", "volume_in", "volume_out",
                                   "mass_in", "mass_out")]


class HgProgress(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("module_time", ctypes.c_double),
                ("dt", ctypes.c_double), ("dt_advection", ctypes.c_double),
                ("steps", ctypes.c_int64),
                ("iteration", ctypes.c_int32),
                ("convergence_indicator", ctypes.c_double), ("meshpos", _d2)]


# hg_output_entry: the module's output entries in writing order
OUTPUT_ENTRIES = ["x", "y", "z", "velocity_x", "velocity_y", "velocity_z", "pressure",
                  "density", "viscosity", "radiation", "volume_source", "temperature",
                  "excluded", "divergence"] + [
    f"{base}{i}" for base in ("mass_source_", "mass_fraction_", "density_",
                              "target_density_", "partial_density_", "volume_fraction_")
    for i in range(HG_MAX_PHASES)]
HG_OUT_COUNT = 32


class HgOutputConfig(ctypes.Structure):
    _fields_ = [("abi_version", ctypes.c_int32),
                ("name", ctypes.c_char * 256), ("plt_title", ctypes.c_char * 256),
                ("filename_field", ctypes.c_char * 512),
                ("filename_scalar", ctypes.c_char * 512),
                ("filename_iter_history", ctypes.c_char * 512),
                ("field_output_format", ctypes.c_char * 32),
                ("no_output", ctypes.c_int32), ("no_mesh_output", ctypes.c_int32),
                ("max_frame_index", ctypes.c_int32),
                ("max_frame_scalar_index", ctypes.c_int32),
                ("output_factor", ctypes.c_int32 * 2), ("T", ctypes.c_double),
                ("iter_history_enable", ctypes.c_int32), ("iter_history_n", ctypes.c_int32),
                ("iter_history_sfixed", ctypes.c_int32),
                ("output", ctypes.c_int32 * HG_OUT_COUNT)]


class HgSolveOpts(ctypes.Structure):
    _fields_ = [("solver", ctypes.c_int32), ("tolerance", ctypes.c_double),
                ("max_iters", ctypes.c_int32),
                ("relaxation_factor", ctypes.c_double),
                ("precision", ctypes.c_int32)]


class HgSolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("final_diff", ctypes.c_double),
                ("rel_residual", ctypes.c_double), ("converged", ctypes.c_int32),
                ("ms", ctypes.c_double)]


class HgTrainOpts(ctypes.Structure):
    _fields_ = [("rhs_count", ctypes.c_int32), ("mg_iters", ctypes.c_int32),
                ("steps", ctypes.c_int32), ("lr", ctypes.c_float),
                ("seed", ctypes.c_uint64), ("flags", ctypes.c_int32)]


class HgTrainStats(ctypes.Structure):
    _fields_ = [("pairs", ctypes.c_int32), ("steps", ctypes.c_int32),
                ("loss_first", ctypes.c_double), ("loss_last", ctypes.c_double),
                ("ms", ctypes.c_double), ("accepted", ctypes.c_int32),
                ("val_iters", ctypes.c_int32), ("val_iters_baseline", ctypes.c_int32),
                ("val_ms", ctypes.c_double), ("val_ms_baseline", ctypes.c_double)]


PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libhydrogpu.so")

# Every entry point include/hydro_gpu.h, include/hydro_output.h and
# include/hydro_noalloc.h declare.
EXPORTS = [
    "hg3_config_default", "hg3_config_parse", "hg3_create", "hg3_destroy", "hg3_last_error",
    "hg3_step", "hg3_get_field", "hg3_set_field",
    "hg_get_config", "hg_get_progress", "hg_get_iteration_history",
    "hg_output_config_default", "hg_output_config_parse", "hg_output_open",
    "hg_output_write_results", "hg_output_close", "hg_output_last_error", "hg_vts_write",
    "hg_config_default", "hg_config_parse", "hg_solver_from_name",
    "hg_create", "hg_destroy", "hg_last_error", "hg_num_cells", "hg_num_faces",
    "hg_set_field", "hg_get_field", "hg_refresh",
    "hg_fluid_start_step", "hg_fluid_iterate", "hg_fluid_is_converged",
    "hg_fluid_finish_step", "hg_advect_substep", "hg_update_properties",
    "hg_calc_stats", "hg_step", "hg_last_step_timing", "hg_time_kernel", "hg_batch_run",
    "hg_stencil_solve", "hg_stencil_solve_noalloc", "hg_poisson_create", "hg_poisson_destroy",
    "hg_poisson_solve", "hg_poisson_time_kernel",
    "hg_train_opts_default", "hg_poisson_train", "hg_train_pressure_net",
    "hg_net_create", "hg_net_destroy", "hg_net_num_weights", "hg_net_configure",
    "hg_net_set_weights",
    "hg_net_get_weights", "hg_net_forward", "hg_net_gradient", "hg_net_train_step",
    "hg_net_time",
]

_P = ctypes.POINTER
_dp = _P(ctypes.c_double)
_fp = _P(ctypes.c_float)


def _declare(lib):
    vp = ctypes.c_void_p
    sig = {
        "hg_config_default": (None, [_P(HgConfig)]),
        "hg3_config_default": (None, [_P(Hg3Config)]),
        "hg3_config_parse": (ctypes.c_int, [_P(Hg3Config), ctypes.c_char_p,
                                            ctypes.c_char_p, ctypes.c_size_t]),
        "hg3_create": (ctypes.c_int, [_P(Hg3Config), _P(vp)]),
        "hg3_destroy": (None, [vp]),
        "hg3_last_error": (ctypes.c_char_p, [vp]),
        "hg3_step": (ctypes.c_int, [vp, _P(HgStepStats)]),
        "hg3_get_field": (ctypes.c_int, [vp, ctypes.c_int32, _dp, ctypes.c_size_t]),
        "hg3_set_field": (ctypes.c_int, [vp, ctypes.c_int32, _dp, ctypes.c_size_t]),
        "hg_config_parse": (ctypes.c_int, [_P(HgConfig), ctypes.c_char_p,
                                           ctypes.c_char_p, ctypes.c_size_t]),
        "hg_solver_from_name": (ctypes.c_int32, [ctypes.c_char_p]),
        "hg_create": (ctypes.c_int, [_P(HgConfig), _P(vp)]),
        "hg_destroy": (None, [vp]),
        "hg_last_error": (ctypes.c_char_p, [vp]),
        "hg_num_cells": (ctypes.c_int32, [vp]),
        "hg_num_faces": (ctypes.c_int32, [vp]),
        "hg_set_field": (ctypes.c_int, [vp, ctypes.c_int32, _dp, ctypes.c_size_t]),
        "hg_get_field": (ctypes.c_int, [vp, ctypes.c_int32, _dp, ctypes.c_size_t]),
        "hg_refresh": (ctypes.c_int, [vp]),
        "hg_get_config": (ctypes.c_int, [vp, _P(HgConfig)]),
        "hg_get_progress": (ctypes.c_int, [vp, _P(HgProgress)]),
        "hg_get_iteration_history": (ctypes.c_int, [vp, _dp, ctypes.c_int32,
                                                    _P(ctypes.c_int32)]),
        "hg_output_config_default": (None, [_P(HgOutputConfig)]),
        "hg_output_config_parse": (ctypes.c_int, [_P(HgOutputConfig), ctypes.c_char_p,
                                                  ctypes.c_char_p, ctypes.c_size_t]),
        "hg_output_open": (ctypes.c_int, [vp, _P(HgOutputConfig), _P(vp)]),
        "hg_output_write_results": (ctypes.c_int, [vp, ctypes.c_int32]),
        "hg_output_close": (ctypes.c_int, [vp]),
        "hg_output_last_error": (ctypes.c_char_p, [vp]),
        "hg_vts_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, _dp, _dp,
                                        ctypes.c_int32, _P(ctypes.c_char_p), _P(_dp),
                                        ctypes.c_int32, _P(ctypes.c_char_p), _P(_dp)]),
        "hg_fluid_start_step": (ctypes.c_int, [vp]),
        "hg_fluid_iterate": (ctypes.c_int, [vp, _P(HgIterStats)]),
        "hg_fluid_is_converged": (ctypes.c_int32, [vp]),
        "hg_fluid_finish_step": (ctypes.c_int, [vp]),
        "hg_advect_substep": (ctypes.c_int, [vp]),
        "hg_update_properties": (ctypes.c_int, [vp]),
        "hg_calc_stats": (ctypes.c_int, [vp, _P(HgStats)]),
        "hg_step": (ctypes.c_int, [vp, _P(HgStepStats)]),
        "hg_last_step_timing": (ctypes.c_int, [vp, _dp, _dp, _dp]),
        "hg_time_kernel": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_int32, _dp]),
        "hg_batch_run": (ctypes.c_int, [_P(vp), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                        _P(HgStepStats), _dp]),
        "hg_stencil_solve": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32] + [_dp] * 7 +
                             [_P(HgSolveOpts), _P(HgSolveStats)]),
        "hg_stencil_solve_noalloc": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32] + [vp] * 7 +
                                     [_P(HgSolveOpts), _P(HgSolveStats), vp]),
        "hg_poisson_create": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _dp, _dp, _P(vp)]),
        "hg_poisson_destroy": (None, [vp]),
        "hg_poisson_solve": (ctypes.c_int, [vp, _dp, _dp, _P(HgSolveOpts), _P(HgSolveStats)]),
        "hg_poisson_time_kernel": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_int32, _dp]),
        "hg_train_opts_default": (None, [_P(HgTrainOpts)]),
        "hg_poisson_train": (ctypes.c_int, [vp, _P(HgTrainOpts), _P(HgTrainStats)]),
        "hg_train_pressure_net": (ctypes.c_int, [vp, _P(HgTrainOpts), _P(HgTrainStats)]),
        "hg_net_create": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _P(vp)]),
        "hg_net_destroy": (None, [vp]),
        "hg_net_num_weights": (ctypes.c_int64, [vp]),
        "hg_net_configure": (ctypes.c_int, [vp, ctypes.c_int32]),
        "hg_net_set_weights": (ctypes.c_int, [vp, _fp]),
        "hg_net_get_weights": (ctypes.c_int, [vp, _fp]),
        "hg_net_forward": (ctypes.c_int, [vp, _fp, _fp]),
        "hg_net_gradient": (ctypes.c_int, [vp, _fp, _fp, _fp]),
        "hg_net_train_step": (ctypes.c_int, [vp, _fp, _fp, ctypes.c_float, _dp]),
        "hg_net_time": (ctypes.c_int, [vp, ctypes.c_int32, _dp, _dp, _dp, _dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


_LIB = None


def load(path=None):
    """Loads libhydrogpu.so; raises if it is missing (no fallback path)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(p)
    _declare(lib)
    if path is None:
        _LIB = lib
    return lib


def default_config(lib=None):
    lib = lib or load()
    cfg = HgConfig()
    lib.hg_config_default(ctypes.byref(cfg))
    return cfg


def parse_config(text, cfg=None, lib=None):
    """Applies .hydroconf `set` lines to cfg (defaults first if cfg is None)."""
    lib = lib or load()
    cfg = cfg if cfg is not None else default_config(lib)
    err = ctypes.create_string_buffer(512)
    rc = lib.hg_config_parse(ctypes.byref(cfg), text.encode(), err, 512)
    if rc != HG_OK:
        raise ValueError(err.value.decode())
    return cfg


def parse_config3(text, lib=None):
    """hg3_config from .hydroconf `set` lines (defaults first)."""
    lib = lib or load()
    cfg = Hg3Config()
    lib.hg3_config_default(ctypes.byref(cfg))
    err = ctypes.create_string_buffer(512)
    rc = lib.hg3_config_parse(ctypes.byref(cfg), text.encode(), err, 512)
    if rc != HG_OK:
        raise ValueError(err.value.decode())
    return cfg


def default_output_config(lib=None):
    lib = lib or load()
    oc = HgOutputConfig()
    lib.hg_output_config_default(ctypes.byref(oc))
    return oc


def parse_output_config(text, oc=None, lib=None):
    """Applies the output keys of .hydroconf `set` lines to oc (defaults
    first if oc is None)."""
    lib = lib or load()
    oc = oc if oc is not None else default_output_config(lib)
    err = ctypes.create_string_buffer(512)
    rc = lib.hg_output_config_parse(ctypes.byref(oc), text.encode(), err, 512)
    if rc != HG_OK:
        raise ValueError(err.value.decode())
    return oc


def vts_write(path, xn, yn, node_arrays, cell_arrays, lib=None):
    """hg_vts_write: node_arrays / cell_arrays are lists of (name, float64
    array) on the (len(xn)-1) x (len(yn)-1) grid."""
    import numpy as np
    lib = lib or load()
    xn = np.ascontiguousarray(xn, dtype=np.float64)
    yn = np.ascontiguousarray(yn, dtype=np.float64)
    keep = []

    def pack(arrays):
        names = (ctypes.c_char_p * max(1, len(arrays)))(*[n.encode() for n, _ in arrays])
        vals = [np.ascontiguousarray(a, dtype=np.float64) for _, a in arrays]
        keep.extend(vals)
        ptrs = (_dp * max(1, len(arrays)))(*[dptr(v) for v in vals])
        return names, ptrs

    nn, npp = pack(node_arrays)
    cn, cp = pack(cell_arrays)
    rc = lib.hg_vts_write(path.encode(), len(xn) - 1, len(yn) - 1, dptr(xn), dptr(yn),
                          len(node_arrays), nn, npp, len(cell_arrays), cn, cp)
    if rc != HG_OK:
        raise RuntimeError(f"hg_vts_write: {STATUS_NAMES.get(rc, rc)}")


def dptr(arr):
    """Pointer to a C-contiguous float64 numpy array."""
    return arr.ctypes.data_as(_dp)


def fptr(arr):
    """Pointer to a C-contiguous float32 numpy array."""
    return arr.ctypes.data_as(_fp)
