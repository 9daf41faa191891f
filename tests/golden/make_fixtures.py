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

    python tests/golden/make_fixtures.py
"""
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle  # noqa: E402
from tests import cases  # noqa: E402

REFERENCE = "/root/reference"
GENERAL = os.path.join(REFERENCE, "examples", "general.hydroconf")
SAMPLE_LOG = os.path.join(REFERENCE, "examples", "cavity", "sample", "exp.log")

FIXTURES = {
    # name: (case text, steps, dump steps)
    "ref_cavity16": (cases.cavity(16), 5, (1, 5)),
    "ref_cavity64": (cases.cavity(64), 20, (20,)),
    "ref_dam64x32": (cases.dam(64, 32), 5, (1, 5)),
    "ref_cavity32_jacobi": (cases.cavity(32, solver_p="jacobi", extra="set double lu_relaxed_relaxation_factor 0.9"), 3, (3,)),
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
    "ref_cavity64_200": (cases.cavity(64), 200, (200,)),
    "ref_dam256x128_50": (cases.dam(256, 128), 50, (50,)),
}


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


def main():
    if not os.path.exists(pyoracle.REF_BIN):
        raise SystemExit("build oracle/_ref first: make -f oracle/ref.mk")
    sample()
    for name, (text, steps, dumps) in FIXTURES.items():
        reference_dump(name, text, steps, dumps)
    for name, (text, steps, dumps) in LONG.items():
        reference_hashes(name, text, steps, dumps)


if __name__ == "__main__":
    main()
