"""Writes the golden fixtures under tests/golden/ (run in the build container,
where /root/reference exists; the fixtures are committed, the reference is
not needed afterwards).

1. cavity_sample.hydroconf / cavity_sample_rs.txt: the parameter dump and the
   1000 per-iteration convergence indicators Rs that the reference ships in
   examples/cavity/sample/exp.log (lines 1-170 and 174-1173): data only.
2. ref_*.npz: FP64 fields dumped by the reference module itself
   (oracle/_ref/hydro_ref, built from the reference sources by oracle/ref.mk
   with -O2 -ffp-contract=off) after given steps, with the SIMPLE iteration
   counts and pressure-solver sweep counts it printed.
3. ref_*.json: the same for long trajectories, stored compactly: the
   SHA-256 of each FP64 field's bytes (bit-for-bit identity) and its L2 norm
   and maximum (to report how far a mismatch is), per dumped step.
4. ref_output_*.npz: every file the reference module's own writers produce
   (ParaView .vts frames and .pvd collection, scalar .dat table, Tecplot
   iteration history) for a case run as the reference console runs it
   (step, advance the time, write_results), with the case and output keys.
5. sample_*.gz: the field frame and iteration history the reference ships
   with its cavity sample run (examples/cavity/sample), gzip-compressed:
   data only.
6. ref3d_*.npz / ref3d_*.json: the same for the reference's 3D module
   (oracle/_ref/hydro_ref3d: hydro<geom3d::MeshStructured<double>>).

    python tests/golden/make_fixtures.py [output | only NAME...]
"""
import os
import re
import gzip
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle  # noqa: E402
from hydro_amd import workloads  # noqa: E402

REFERENCE = "/root/reference"
GENERAL = os.path.join(REFERENCE, "examples", "general.hydroconf")
SAMPLE_LOG = os.path.join(REFERENCE, "examples", "cavity", "sample", "exp.log")

DT_AUTO = workloads.DT_AUTO

FIXTURES = {
    # name: (case text, steps, dump steps)
    "ref_cavity16": (workloads.cavity(16), 5, (1, 5)),
    "ref_cavity64": (workloads.cavity(64), 20, (20,)),
    "ref_dam64x32": (workloads.dam(64, 32), 5, (1, 5)),
    "ref_cavity32_jacobi": (workloads.cavity(32, solver_p="jacobi", extra="set double lu_relaxed_relaxation_factor 0.9"), 3, (3,)),
    "ref_dam32x16_dtauto": (workloads.dam(32, 16, extra=DT_AUTO), 6, (1, 6)),
    "ref_cavity32_heat": (workloads.cavity(32, extra=workloads.HEAT_BOTTOM), 5, (1, 5)),
    "ref_dam32x16_heat": (workloads.dam(32, 16, extra=workloads.HEAT_LEFT), 4, (4,)),
    "ref_cavity32_simpler": (workloads.cavity(32, extra="set bool simpler 1"), 5, (1, 5)),
    "ref_dam32x16_simpler": (workloads.dam(32, 16, extra="set bool simpler 1"), 4, (4,)),
    "ref_channel32x16": (workloads.channel(32, 16), 6, (1, 6)),
    "ref_channel32x16_jacobi": (workloads.channel(32, 16, extra=(
        "set string linear_solver_pressure jacobi\n"
        "set double lu_relaxed_relaxation_factor 0.9\n")), 4, (4,)),
    "ref_cavity32_simpler_pinned": (workloads.cavity(32, solver_p="jacobi", extra=(
        "set bool simpler 1\nset double lu_relaxed_relaxation_factor 0.9\n"
        "set vect pressure_fixed_point (0.3, 0.6)\nset double pressure_fixed_value 0.1\n")),
        4, (4,)),
    # LuDecompositionRelaxed (linear.hpp:578-651) for the pressure: at the
    # default tolerance one pass per solve; at 1e-5 every solve runs to the
    # 1000-iteration limit (the backward step reads the previous iterate)
    "ref_cavity32_lu_relaxed": (workloads.cavity(32, solver_p="lu_relaxed"), 3, (1, 3)),
    "ref_cavity32_lu_relaxed_tight": (workloads.cavity(32, solver_p="lu_relaxed", extra=(
        "set double lu_relaxed_tolerance 1e-5\n")), 2, (2,)),
    # lu_relaxed and jacobi as the velocity solver (conv_diff.hpp:245): each
    # SIMPLE iteration prints the two velocity solves before the pressure's
    "ref_cavity32_vel_lu_relaxed": (workloads.cavity(32, solver_p="jacobi", solver_v="lu_relaxed", extra=(
        "set double lu_relaxed_tolerance 1e-6\nset double lu_relaxed_relaxation_factor 0.5\n")),
        3, (1, 3)),
    "ref_dam32x16_vel_jacobi": (workloads.dam(32, 16, solver_p="lu_relaxed", solver_v="jacobi", extra=(
        "set double lu_relaxed_tolerance 1e-5\nset double lu_relaxed_relaxation_factor 0.9\n")),
        3, (1, 3)),
}


def sample():
    lines = open(SAMPLE_LOG).read().splitlines()
    params = [l for l in lines if l.startswith("set ")]
    rs = [m.group(1) for l in lines for m in [re.match(r"\.\.\.\.\.s=\d+, Rs=(\S+)", l)] if m]
    with open(os.path.join(HERE, "cavity_sample.hydroconf"), "w") as f:
        f.write("# parameter dump of examples/cavity/sample/exp.log (reference data)\n")
        f.write("\n".join(params) + "\n")
    with open(os.path.join(HERE, "cavity_sample_rs.txt"), "w") as f:
        f.write("# Rs per SIMPLE iteration, examples/cavity/sample/exp.log:174-1173\n")
        f.write("\n".join(rs) + "\n")
    print(f"sample: {len(params)} parameters, {len(rs)} Rs values")


def reference_dump(name, text, steps, dumps):
    script = open(GENERAL).read() + "\n" + text + "\nset bool no_output 1\n"
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "case.conf")
        open(path, "w").write(script)
        r = subprocess.run([pyoracle.REF_BIN, path, str(steps), os.path.join(tmp, "d"), "1"],
                           capture_output=True, text=True, cwd=tmp)
        if r.returncode != 0:
            raise RuntimeError(r.stderr)
        out = {"text": np.array(text), "steps": np.array(steps)}
        iters = [int(m.group(1)) for m in re.finditer(r"STEP s=\d+ iters=(\d+)", r.stdout)]
        sweeps = [int(m.group(1)) for m in re.finditer(r"iter = (\d+), diff", r.stdout)]
        out["simple_iters"] = np.array(iters)
        out["pressure_sweeps"] = np.array(sweeps)
        for s in dumps:
            d = pyoracle.read_ref_dump(os.path.join(tmp, f"d.{s}.bin"))
            for k, v in d.items():
                if isinstance(v, np.ndarray):
                    out[f"s{s}_{k}"] = v
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print(f"{name}: {steps} steps, SIMPLE iterations {iters}")


LONG = {
    # C1: lid-driven cavity 64x64, 200 steps; C2: broken dam 256x128, the
    # first 50 steps (later steps of this case need the 1000-sweep limit of
    # the pressure solver on every SIMPLE iteration)
    "ref_cavity64_200": (workloads.cavity(64), 200, (200,)),
    "ref_dam256x128_50": (workloads.dam(256, 128), 50, (50,)),
    # C2 in full: 500 steps (about 40 minutes of the reference module; from
    # about step 200 its SOR runs to the 1000-sweep limit every iteration)
    "ref_dam256x128_500": (workloads.dam(256, 128), 500, (100, 200, 300, 400, 500)),
}


FIXTURES3 = {
    # cavity with the jacobi solvers, one SIMPLE iteration per step
    "ref3d_cavity12x10x8": (workloads.cavity3d(12, 10, 8), 5, (1, 5)),
    # the SIMPLE loop to convergence (3-4 iterations per step)
    "ref3d_cavity12x10x8_loop": (workloads.cavity3d(12, 10, 8, extra=(
        "set double convergence_tolerance 1e-4\n")), 4, (4,)),
    # the reference's default solvers (gauss_seidel pressure, lu velocity)
    "ref3d_cavity10x8x6_gs_lu": (workloads.cavity3d(10, 8, 6, solver_p="gauss_seidel",
                                                      solver_v="lu"), 3, (3,)),
    # two phases, gravity in z
    "ref3d_dam16x4x8": (workloads.dam3d(16, 4, 8, dt=1e-3), 4, (1, 4)),
}
LONG3 = {
    # cavity 32^3, 60 steps
    "ref3d_cavity32_60": (workloads.cavity3d(32, 32, 32), 60, (20, 60)),
}


def _run3(text, steps, every, tmp):
    script = open(GENERAL).read() + "\n" + text + "\nset bool no_output 1\n"
    path = os.path.join(tmp, "case.conf")
    open(path, "w").write(script)
    r = subprocess.run([pyoracle.REF_BIN_3D, path, str(steps), os.path.join(tmp, "d"),
                        str(every)], capture_output=True, text=True, cwd=tmp)
    if r.returncode != 0:
        raise RuntimeError(r.stderr)
    return r.stdout


def reference_dump3(name, text, steps, dumps):
    with tempfile.TemporaryDirectory() as tmp:
        stdout = _run3(text, steps, 1, tmp)
        out = {"text": np.array(text), "steps": np.array(steps)}
        out["simple_iters"] = np.array([int(m.group(1)) for m in
                                        re.finditer(r"STEP s=\d+ iters=(\d+)", stdout)])
        out["pressure_sweeps"] = np.array([int(m.group(1)) for m in
                                           re.finditer(r"iter = (\d+), diff", stdout)])
        for s in dumps:
            d = pyoracle.read_ref_dump3(os.path.join(tmp, f"d.{s}.bin"))
            for k, v in d.items():
                if isinstance(v, np.ndarray):
                    out[f"s{s}_{k}"] = v
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print(f"{name}: {steps} steps, SIMPLE iterations {list(out['simple_iters'])}")


def reference_hashes3(name, text, steps, dumps):
    import hashlib
    import json
    with tempfile.TemporaryDirectory() as tmp:
        stdout = _run3(text, steps, min(dumps), tmp)
        out = {"text": text, "steps": steps,
               "simple_iters": [int(m.group(1)) for m in
                                re.finditer(r"STEP s=\d+ iters=(\d+)", stdout)],
               "dumps": {}}
        for s in dumps:
            d = pyoracle.read_ref_dump3(os.path.join(tmp, f"d.{s}.bin"))
            out["dumps"][str(s)] = {
                k: {"sha256": hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest(),
                    "l2": float(np.linalg.norm(v)), "max": float(np.abs(v).max())}
                for k, v in d.items() if isinstance(v, np.ndarray)}
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(out, f, indent=1)
        print(f"{name}: {steps} steps, SIMPLE iterations total {sum(out['simple_iters'])}")


def reference_hashes(name, text, steps, dumps):
    import hashlib
    import json
    script = open(GENERAL).read() + "\n" + text + "\nset bool no_output 1\n"
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "case.conf")
        open(path, "w").write(script)
        r = subprocess.run([pyoracle.REF_BIN, path, str(steps), os.path.join(tmp, "d"),
                            str(min(dumps))], capture_output=True, text=True, cwd=tmp)
        if r.returncode != 0:
            raise RuntimeError(r.stderr)
        out = {"text": text, "steps": steps,
               "simple_iters": [int(m.group(1)) for m in
                                re.finditer(r"STEP s=\d+ iters=(\d+)", r.stdout)],
               "pressure_sweeps_total": sum(int(m.group(1)) for m in
                                            re.finditer(r"iter = (\d+), diff", r.stdout)),
               "dumps": {}}
        for s in dumps:
            d = pyoracle.read_ref_dump(os.path.join(tmp, f"d.{s}.bin"))
            out["dumps"][str(s)] = {
                k: {"sha256": hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest(),
                    "l2": float(np.linalg.norm(v)), "max": float(np.abs(v).max())}
                for k, v in d.items() if isinstance(v, np.ndarray)}
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(out, f, indent=1)
        print(f"{name}: {steps} steps, SIMPLE iterations total {sum(out['simple_iters'])}")


# Output keys of the output fixtures: several frames and scalar rows, every
# entry the GPU path can supply, an iteration history.
DAM_OUTPUT = """
set string name exp
set string plt_title Dam_32x16
set double T 0.0008
set int max_frame_index 3
set int max_frame_scalar_index 5
set int output_factor_x 2
set bool output_z 1
set bool output_velocity_z 1
set bool output_density 1
set bool output_viscosity 1
set bool output_radiation 1
set bool output_excluded 1
set bool output_divergence 1
set bool output_mass_fraction_0 1
set bool output_mass_fraction_1 1
set bool output_density_0 1
set bool output_density_1 1
set bool output_target_density_0 1
set bool output_target_density_1 1
set bool output_partial_density_0 1
set bool output_partial_density_1 1
set bool iter_history_enable 1
set int iter_history_n 2
"""
CAVITY_MESHVEL_OUTPUT = """
set string name cav
set string plt_title Cavity_16_meshvel
set double T 0.03
set int max_frame_index 2
set int max_frame_scalar_index 10
set int output_factor_y 2
set bool output_divergence 1
set bool iter_history_enable 1
set int iter_history_n 0
"""
DTAUTO_OUTPUT = """
set string name dta
set double T 0.02
set int max_frame_index 2
set int max_frame_scalar_index 40
set bool iter_history_enable 1
set int iter_history_n 1
"""
HEAT_OUTPUT = """
set string name heat
set double T 0.0006
set int max_frame_index 2
set int max_frame_scalar_index 3
set bool output_temperature 1
"""
CHANNEL_OUTPUT = """
set string name chan
set double T 0.02
set int max_frame_index 2
set int max_frame_scalar_index 5
set bool output_divergence 1
"""
OUTPUT = {
    # name: (case text, output keys, step limit; the run ends at T)
    "ref_output_dam32x16": (workloads.dam(32, 16), DAM_OUTPUT, 100),
    "ref_output_cavity16_meshvel": (
        workloads.cavity(16, extra="set vect meshvel (0.3, -0.2)"), CAVITY_MESHVEL_OUTPUT, 100),
    "ref_output_dam32x16_dtauto": (workloads.dam(32, 16, extra=DT_AUTO), DTAUTO_OUTPUT, 100),
    "ref_output_dam32x16_heat": (workloads.dam(32, 16, extra=workloads.HEAT_BOTTOM), HEAT_OUTPUT, 100),
    "ref_output_channel32x16": (workloads.channel(32, 16), CHANNEL_OUTPUT, 100),
}


def reference_output(name, text, out_text, steps):
    script = open(GENERAL).read() + "\n" + text + "\n" + out_text
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "case.conf")
        open(path, "w").write(script)
        r = subprocess.run([pyoracle.REF_BIN, path, str(steps), "-", "-1"],
                           capture_output=True, text=True, cwd=tmp)
        if r.returncode != 0:
            raise RuntimeError(r.stderr)
        files = sorted(f for f in os.listdir(tmp) if f != "case.conf")
        contents = [open(os.path.join(tmp, f)).read() for f in files]
        nsteps = len(re.findall(r"STEP s=", r.stdout))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), text=np.array(text),
                            output=np.array(out_text), steps=np.array(nsteps),
                            files=np.array(files), contents=np.array(contents))
        print(f"{name}: {nsteps} steps, files {files}")


def sample_files():
    for f in ("exp.field.0.vts", "exp.iter_history.plt"):
        data = open(os.path.join(os.path.dirname(SAMPLE_LOG), f), "rb").read()
        with open(os.path.join(HERE, "sample_" + f + ".gz"), "wb") as raw, \
                gzip.GzipFile(fileobj=raw, mode="wb", mtime=0) as g:
            g.write(data)
        print(f"sample_{f}.gz: {len(data)} bytes")


def main():
    """No argument: every fixture.  `output`: the output files and samples;
    `only NAME...`: the named fixtures."""
    if not os.path.exists(pyoracle.REF_BIN):
        raise SystemExit("build oracle/_ref first: make -f oracle/ref.mk")
    args = sys.argv[1:]
    if args[:1] == ["only"]:
        for name in args[1:]:
            if name in FIXTURES:
                reference_dump(name, *FIXTURES[name])
            elif name in OUTPUT:
                reference_output(name, *OUTPUT[name])
            elif name in LONG:
                reference_hashes(name, *LONG[name])
            elif name in FIXTURES3:
                reference_dump3(name, *FIXTURES3[name])
            elif name in LONG3:
                reference_hashes3(name, *LONG3[name])
            else:
                raise SystemExit(f"unknown fixture {name}")
        return
    sample_files()
    for name, (text, out_text, steps) in OUTPUT.items():
        reference_output(name, text, out_text, steps)
    if args == ["output"]:
        return
    sample()
    for name, (text, steps, dumps) in FIXTURES.items():
        reference_dump(name, text, steps, dumps)
    for name, (text, steps, dumps) in LONG.items():
        reference_hashes(name, text, steps, dumps)
    for name, (text, steps, dumps) in FIXTURES3.items():
        reference_dump3(name, text, steps, dumps)
    for name, (text, steps, dumps) in LONG3.items():
        reference_hashes3(name, text, steps, dumps)


if __name__ == "__main__":
    main()
