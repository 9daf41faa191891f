"""The drop-in at the solve level, compiled against the reference's classes.

oracle/_ref/hydro_ref_gpu is the reference module's driver built from the
reference sources (oracle/ref.mk) with the reference's own FluidSimple
(fluid.hpp:422-1169) whose pressure LinearSolver -- and, in mode "all", the
velocity components' solvers -- are GpuLinearSolver adapters over
hg_stencil_solve in libhydrogpu.so (oracle/ref_gpu_dispatch.cpp).  The run
must equal the reference module run alone (oracle/_ref/hydro_ref, same
driver, the reference's CPU solvers) bit for bit: every dumped field, the
SIMPLE iterations of every step, and the iteration count and final
difference every solver prints.
"""
import os
import subprocess

import numpy as np
import pytest

from hydro_amd import workloads
from oracle import pyoracle

pytestmark = pytest.mark.gpu

REF_DIR = os.path.dirname(pyoracle.REF_BIN)
REF_GPU_BIN = os.path.join(REF_DIR, "hydro_ref_gpu")
GENERAL = os.path.join(REF_DIR, "general.hydroconf")

CASES = {
    "cavity32_gs_lu": workloads.cavity(32),
    "dam32x16_lurelaxed_jacobi": workloads.dam(32, 16, solver_p="lu_relaxed", solver_v="jacobi",
                                               extra="set double lu_relaxed_tolerance 1e-5\n"
                                                     "set double lu_relaxed_relaxation_factor 0.9\n"),
}


def run(exe, script, steps, prefix, mode=None):
    args = [exe, script, str(steps), prefix, str(steps)] + ([mode] if mode else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (exe, r.returncode, r.stderr[-2000:])
    lines = [l for l in r.stdout.splitlines() if l.startswith("iter =") or l.startswith("STEP")]
    # the step timing differs between runs: keep the counts only
    return [l.split(" t=")[0] for l in lines]


@pytest.mark.parametrize("case,mode", [("cavity32_gs_lu", "pressure"), ("cavity32_gs_lu", "all"),
                                       ("dam32x16_lurelaxed_jacobi", "all")])
def test_reference_fluidsimple_on_gpu_solvers(gpu, tmp_path, case, mode):
    for exe in (pyoracle.REF_BIN, REF_GPU_BIN, GENERAL):
        if not os.path.exists(exe):
            pytest.skip(f"{exe} not built (oracle/ref.mk needs the reference sources)")
    steps = 5
    print(f"case: {case}, {steps} steps, reference solvers vs GPU dispatch ({mode})", flush=True)
    script = tmp_path / "case.conf"
    script.write_text(open(GENERAL).read() + CASES[case] + "\nset bool no_output 1\n")
    cpu = run(pyoracle.REF_BIN, str(script), steps, str(tmp_path / "cpu"))
    gpu_ = run(REF_GPU_BIN, str(script), steps, str(tmp_path / "gpu"), mode)
    assert cpu == gpu_
    a = pyoracle.read_ref_dump(str(tmp_path / f"cpu.{steps}.bin"))
    b = pyoracle.read_ref_dump(str(tmp_path / f"gpu.{steps}.bin"))
    for k, v in a.items():
        if isinstance(v, np.ndarray):
            assert np.array_equal(v, b[k]), k
