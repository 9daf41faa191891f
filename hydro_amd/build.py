"""Build recipes: libhydrogpu.so (HIP, gfx950), the CPU oracle and the
reference build used as the CPU checker / baseline.

    python -m hydro_amd.build          (or __graft_entry__.build())

Everything is built in-tree so the shared objects travel with the tree.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "hydro_amd")
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libhydrogpu.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")
REFERENCE = "/root/reference"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = "gfx950"

# Parity-path sources evaluate every expression in the reference's order:
# no contraction into FMAs.  The throughput solvers and the network allow it.
EXACT = {"step_kernels.hip", "solvers.hip", "context.hip", "config.cpp", "output.cpp", "step3d.hip"}
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument", "-I", os.path.join(ROOT, "include")]


def _run(cmd, cwd=ROOT):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def _compile(src):
    name = os.path.basename(src)
    obj = os.path.join(BUILD, name + ".o")
    flags = list(COMMON)
    flags.append("-ffp-contract=off" if name in EXACT else "-ffp-contract=fast")
    if name.endswith(".cpp"):
        flags += ["-x", "hip"]
    _run([HIPCC] + flags + ["-c", src, "-o", obj])
    return obj


def build_gpu_lib():
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


MICRO = os.path.join(ROOT, "hydro_amd", "micro")


def build_micro():
    """The launch and barrier microbenchmarks (standalone programs, not
    part of the library), next to their sources."""
    out = []
    for name, extra in (("grid_barrier", []), ("graph_launch", ["-pthread"])):
        exe = os.path.join(MICRO, name)
        _run([HIPCC, "-O3", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
              "-packed-fp32-ops", "-Wno-unused-command-line-argument",
              os.path.join(MICRO, name + ".hip"), "-o", exe] + extra)
        out.append(exe)
    return out


def build_oracle():
    tmp = ORACLE_LIB + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared",
          "-o", tmp, os.path.join(ORACLE_DIR, "hydro_oracle.cpp"),
          os.path.join(ORACLE_DIR, "hydro_oracle3d.cpp")])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build_reference():
    """Compiles the reference module (oracle/ref.mk) when its sources exist."""
    if not os.path.isdir(os.path.join(REFERENCE, "source")):
        return None
    _run(["make", "-s", "-f", "oracle/ref.mk", "-j4", f"REF={REFERENCE}", "all"])
    _run(["make", "-s", "-f", "oracle/ref.mk", f"REF={REFERENCE}", "ref3d"])
    if os.path.exists(LIB):
        _run(["make", "-s", "-f", "oracle/ref.mk", f"REF={REFERENCE}", "gpu"])
    return os.path.join(ORACLE_DIR, "_ref")


def build_all(verbose=True):
    out = {}
    out["oracle"] = build_oracle()
    out["gpu"] = build_gpu_lib()
    out["reference"] = build_reference()
    out["micro"] = build_micro()
    if verbose:
        for k, v in out.items():
            print(f"[build] {k}: {v}", flush=True)
    return out


if __name__ == "__main__":
    build_all()
    sys.exit(0)
