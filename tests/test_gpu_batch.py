"""Many cases at once (hg_batch_run): every case stepped concurrently on its
own stream gives bit for bit what it gives alone, with the same SIMPLE
iteration counts."""
import numpy as np
import pytest

from hydro_amd import capi, workloads
from hydro_amd.hydro import Hydro, batch_run

pytestmark = pytest.mark.gpu

FIELDS = [capi.HG_FIELD_VELOCITY_X, capi.HG_FIELD_VELOCITY_Y, capi.HG_FIELD_PRESSURE,
          capi.HG_FIELD_VOLUME_FLUX]


def case_texts():
    return [workloads.cavity(32), workloads.dam(64, 32), workloads.cavity(48, solver_p="jacobi",
                                                                extra="set double lu_relaxed_relaxation_factor 0.9"),
            workloads.dam(32, 16, solver_p="fcg_mg", solver_v="lu_rb", extra="set string precision_gpu mixed"),
            ] + workloads.c5_cases(count=4, seed=7, sides=(64, 96))


def run_alone(text, steps):
    with Hydro(text) as h:
        for _ in range(steps):
            st = h.step()
        return [h.get(f) for f in FIELDS], st.simple_iterations


def test_batch_matches_cases_run_alone(gpu):
    texts = case_texts()
    steps = 3
    print(f"case: batch of {len(texts)}, {steps} steps", flush=True)
    hs = [Hydro(t) for t in texts]
    stats, ms = batch_run(hs, steps, threads=0)
    print(f"batch wall {ms:.1f} ms", flush=True)
    for k, (h, t) in enumerate(zip(hs, texts)):
        ref, iters = run_alone(t, steps)
        for f, r in zip(FIELDS, ref):
            assert np.array_equal(h.get(f), r), (k, f)
        assert stats[k].simple_iterations == iters, k
        h.close()


def test_batch_reports_errors(gpu):
    with pytest.raises(Exception):
        batch_run([], -1)
