"""C5: the batch of 64 independent cases, each one checked.

workloads.c5_cases(64): lid-driven cavities (n x n) and broken dams
(n x n/2) with sides 64..512 and parameters jittered by +-10 %, pressure by
FCG + V-cycle, velocity by lu_rb, the mixed-precision step.  All 64 run at
once through hg_batch_run (the console's concurrent experiment threads,
console.cpp:162-178), one step per call so that every step's statistics are
seen.  For every case:

* every pressure solve of every step reached the true FP64 relative
  residual 1e-8 (hg_step_stats.pressure_converged, pressure_solve_diff_max);
* each phase's volume is conserved to FP64 rounding;
* the corrected volume flux is divergence-free to the solver tolerance:
  ||div flux||_2 <= 2e-8 ||b||_2, b the last pressure system's right-hand
  side (the discrete divergence of the predicted flux, fluid.hpp:946-1027);
* cases with sides <= 128 agree with the CPU oracle (FP64, Jacobi-PCG to the
  same tolerance) within the mixed path's bounds;
* the eight largest cases, stepped again alone, give bit for bit the fields
  and per-step pressure iteration counts of the batch (concurrent contexts
  share nothing, and no kernel's result depends on workgroup timing).
"""
import re

import numpy as np
import pytest

from hydro_amd import capi, compare, workloads
from hydro_amd.hydro import Hydro, batch_run
from oracle import pyoracle

pytestmark = pytest.mark.gpu

STEPS = 3


def divergence(flux, nx, ny):
    nfx = (nx + 1) * ny
    fx = flux[:nfx].reshape(ny, nx + 1)
    fy = flux[nfx:].reshape(ny + 1, nx)
    return ((fx[:, 1:] - fx[:, :-1]) + (fy[1:, :] - fy[:-1, :])).reshape(-1)


def test_c5_batch_of_64_cases(gpu):
    texts = workloads.c5_cases(count=64)
    sizes = [(int(re.search(r"Nx (\d+)", t).group(1)), int(re.search(r"Ny (\d+)", t).group(1)))
             for t in texts]
    print(f"case: C5, {len(texts)} cases, {sum(a * b for a, b in sizes)} cells, {STEPS} steps",
          flush=True)
    hs = [Hydro(t) for t in texts]
    v0 = [compare.phase_volumes(h) for h in hs]
    worst_res = 0.0
    piters = [[] for _ in hs]
    for s in range(STEPS):
        stats, ms = batch_run(hs, 1, threads=8)
        for i, st in enumerate(stats):
            piters[i].append(st.pressure_solve_iters_total)
            assert st.simple_iterations >= 1, i
            assert st.pressure_converged == 1, (i, s, st.pressure_solve_diff_max)
            assert st.pressure_solve_diff_max <= 1e-8, (i, s, st.pressure_solve_diff_max)
            worst_res = max(worst_res, st.pressure_solve_diff_max)
        print(f"  step {s + 1}: {ms:.1f} ms for the batch", flush=True)
    worst_div = worst_drift = 0.0
    for i, h in enumerate(hs):
        nx, ny = sizes[i]
        drift = max(compare.volume_drift(v0[i], compare.phase_volumes(h)))
        assert drift <= 1e-13, (i, drift)
        div = divergence(h.get(capi.HG_FIELD_VOLUME_FLUX), nx, ny)
        b = h.get(capi.HG_FIELD_PRESSURE_RHS)
        rel = np.linalg.norm(div) / np.linalg.norm(b)
        assert rel <= 2e-8, (i, rel)
        worst_div, worst_drift = max(worst_div, rel), max(worst_drift, drift)
    print(f"  worst: solve residual {worst_res:.2e}, flux divergence {worst_div:.2e}, "
          f"volume drift {worst_drift:.2e}", flush=True)
    # the small cases against the CPU oracle
    worst = {}
    n_small = 0
    for i, t in enumerate(texts):
        if max(sizes[i]) > 128:
            continue
        n_small += 1
        o = pyoracle.Oracle(capi.parse_config(t))
        for _ in range(STEPS):
            o.step()
        errs = compare.field_errors(compare.snapshot(hs[i]), compare.snapshot(o))
        for k, v in errs.items():
            worst[k] = max(worst.get(k, 0.0), v)
    print(f"  {n_small} cases <= 128 against the oracle, worst: " +
          ", ".join(f"{k} {v:.2e}" for k, v in worst.items()), flush=True)
    assert n_small >= 1
    bounds = {"u": 1e-5, "v": 1e-5, "p": 1e-6, "alpha0": 1e-9, "alpha1": 1e-9}
    for k, v in worst.items():
        assert v <= bounds[k], (k, v)
    # the largest cases alone: the batch's result, bit for bit
    largest = sorted(range(len(texts)), key=lambda i: -sizes[i][0] * sizes[i][1])[:8]
    fields = [capi.HG_FIELD_VELOCITY_X, capi.HG_FIELD_VELOCITY_Y, capi.HG_FIELD_PRESSURE,
              capi.HG_FIELD_VOLUME_FLUX]
    for i in largest:
        with Hydro(texts[i]) as h1:
            alone = [h1.step().pressure_solve_iters_total for _ in range(STEPS)]
            assert alone == piters[i], (i, sizes[i], alone, piters[i])
            for f in fields:
                assert np.array_equal(h1.get(f), hs[i].get(f)), (i, sizes[i], f)
    print(f"  {len(largest)} largest cases alone: bitwise equal, pressure iterations "
          f"{[piters[i] for i in largest]}", flush=True)
    for h in hs:
        h.close()
