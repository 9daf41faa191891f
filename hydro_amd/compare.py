"""Per-field comparison of two runs of the same case (host-side numpy).

The mixed-precision path is stated against the FP64 path by the relative L2
difference of each field after the same steps from the same state:

* u, v: velocity components;
* p: pressure, each field's mean removed first (the face-form pressure
  solvers fix the pressure only up to a constant: the singular Neumann
  system, fluid.hpp:946-1027 without a fixed pressure point);
* alpha_i: volume fraction of each phase;

and by the drift of the mass check sum_c alpha_i V over the run, for each
path (CalcStat's volume, hydro2d.hpp:1432-1467), in FP64.
"""
import numpy as np

from . import capi


def snapshot(h):
    """The fields compared, from a Hydro or an Oracle, as host arrays."""
    out = {"u": h.get(capi.HG_FIELD_VELOCITY_X), "v": h.get(capi.HG_FIELD_VELOCITY_Y),
           "p": h.get(capi.HG_FIELD_PRESSURE)}
    for i in range(h.cfg.num_phases):
        out[f"alpha{i}"] = h.get(capi.HG_FIELD_VOLUME_FRACTION_0 + i)
    return out


def phase_volumes(h):
    """sum_c alpha_i V per phase (FP64 reduction of the context)."""
    st = h.stats()
    return [float(st.volume[i]) for i in range(h.cfg.num_phases)]


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))


def field_errors(test, ref):
    """Relative L2 difference of every field of `test` against `ref`
    (snapshots); the pressure with each mean removed."""
    out = {}
    for k in ref:
        a, b = test[k], ref[k]
        if k == "p":
            a, b = a - a.mean(), b - b.mean()
        out[k] = rel_l2(a, b)
    return out


def volume_drift(v0, v1):
    """Relative change of each phase's volume sum over a run."""
    return [abs(b - a) / abs(a) if a != 0 else abs(b - a) for a, b in zip(v0, v1)]
