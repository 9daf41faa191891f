"""What hg_step promises around its synchronisation.

* hg_step returns once the SIMPLE loop has finished (its NaN check and the
  step's statistics settled) and leaves the advection, property and stats
  kernels queued on the context's stream; every readback waits for them.
  Stepping one context and reading its fields after each step must give
  the fields of the same steps taken through the fine-grained entry points,
  which synchronise where the reference's calls return.
* The NaN check (FluidSimple::FinishStep, fluid.hpp:1164-1166, and the
  check of the initial fields in StartStep) raises HG_ERR_NAN and leaves the
  device flag clear: after the bad field is replaced the context steps
  normally.
* hg_last_step_timing waits for the step's last kernels and reports the
  three regions.
"""
import math

import numpy as np
import pytest

from hydro_amd import capi
from hydro_amd.hydro import Hydro, HydroNaN
from hydro_amd import workloads

pytestmark = pytest.mark.gpu

FIELDS = (capi.HG_FIELD_VELOCITY_X, capi.HG_FIELD_VELOCITY_Y, capi.HG_FIELD_PRESSURE,
          capi.HG_FIELD_VOLUME_FLUX, capi.HG_FIELD_VOLUME_FRACTION_0)


def test_step_readbacks_match_fine_grained_steps(gpu):
    text = workloads.dam(64, 32, dt=1e-4, solver_p="fcg_mg", solver_v="lu_rb",
                     extra="set string precision_gpu mixed\n")
    print("case: dam 64x32, 4 steps, hg_step vs fine-grained entry points", flush=True)
    with Hydro(text) as a, Hydro(text) as b:
        for _ in range(4):
            sa = a.step()
            b.start_step()
            while not b.is_converged():
                b.iterate()
            b.finish_step()
            for _ in range(sa.advection_substeps):
                b.advect_substep()
            b.update_properties()
            for f in FIELDS:
                assert np.array_equal(a.get(f), b.get(f)), f
        t = a.timing()
        assert t["step.fluid"] > 0 and t["step.advection"] > 0
        assert t["step.fluid_properties"] > 0


def test_nan_raises_and_the_flag_clears(gpu):
    print("case: cavity 16x16 with a NaN pressure cell", flush=True)
    with Hydro(workloads.cavity(16)) as h:
        good = h.get(capi.HG_FIELD_PRESSURE)
        bad = good.copy()
        bad[37] = math.nan
        h.set(capi.HG_FIELD_PRESSURE, bad)
        with pytest.raises(HydroNaN, match="NaN"):
            h.step()
        h.set(capi.HG_FIELD_PRESSURE, good)
        st = h.step()
        assert st.simple_iterations >= 1
        assert np.isfinite(h.get(capi.HG_FIELD_PRESSURE)).all()


def test_nan_velocity_is_named_and_time_does_not_advance(gpu):
    """A NaN velocity is "NaN initial field" (the velocity solver's
    StartStep, conv_diff.hpp:119-121), after the pressure passed; a failed
    step leaves the time and the step count where they were (the checks
    precede IncTime, fluid.hpp:1164-1168)."""
    print("case: cavity 16x16 with a NaN velocity cell", flush=True)
    with Hydro(workloads.cavity(16)) as h:
        h.step()
        before = h.progress()
        good = h.get(capi.HG_FIELD_VELOCITY_X)
        bad = good.copy()
        bad[11] = math.nan
        h.set(capi.HG_FIELD_VELOCITY_X, bad)
        with pytest.raises(HydroNaN, match="NaN initial field"):
            h.step()
        after = h.progress()
        assert (after.time, after.steps) == (before.time, before.steps)
        h.set(capi.HG_FIELD_VELOCITY_X, good)
        h.step()
        assert h.progress().steps == before.steps + 1


def test_time_kernel_leaves_step_timing_intact(gpu):
    """hg_time_kernel times on events of its own: the step regions that
    hg_last_step_timing reports afterwards are still the last step's."""
    print("case: dam 32x16, hg_time_kernel between hg_step and its timing", flush=True)
    with Hydro(workloads.dam(32, 16)) as h:
        h.step()
        t0 = h.timing()
        h.time_kernel("KAdvStep", reps=3)
        t1 = h.timing()
        assert t0 == t1
        assert all(v > 0 for v in t1.values())
