"""The 3D oracle (oracle/hydro_oracle3d.cpp) against the reference's own 3D
module: hydro<geom3d::MeshStructured<double>> (hydro3d.cpp:17-20), built
from the reference sources by oracle/ref.mk (target ref3d) and dumped into
tests/golden/ref3d_*.  Every field of every dumped step is compared bit for
bit, with the SIMPLE iteration count of every step:

* cavity 12x10x8 with the jacobi solvers (one SIMPLE iteration per step);
* the same with the SIMPLE loop run to convergence (3 iterations per step);
* cavity 10x8x6 with the reference's defaults, gauss_seidel pressure and
  lu velocity;
* a two-phase broken dam 16x4x8 with gravity in z (2-6 iterations per step);
* cavity 32^3, 60 steps, against the SHA-256 of the reference's fields.

Plus the 3D configuration reader.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from hydro_amd import capi, workloads
from oracle import pyoracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = {"u": capi.HG_FIELD_VELOCITY_X, "v": capi.HG_FIELD_VELOCITY_Y,
          "w": capi.HG_FIELD_VELOCITY_Z, "p": capi.HG_FIELD_PRESSURE,
          "flux": capi.HG_FIELD_VOLUME_FLUX, "pd0": capi.HG_FIELD_PARTIAL_DENSITY_0,
          "pd1": capi.HG_FIELD_PARTIAL_DENSITY_0 + 1,
          "alpha0": capi.HG_FIELD_VOLUME_FRACTION_0, "alpha1": capi.HG_FIELD_VOLUME_FRACTION_0 + 1}
CASES = ["ref3d_cavity12x10x8", "ref3d_cavity12x10x8_loop", "ref3d_cavity10x8x6_gs_lu",
         "ref3d_dam16x4x8"]


def load(name):
    return np.load(os.path.join(HERE, name + ".npz"))


def dumped_steps(fx):
    return sorted({int(k.split("_")[0][1:]) for k in fx.files if k.startswith("s")
                   and "_" in k and k[1].isdigit()})


def check_case(name, make):
    """Steps `make(cfg3)` (the oracle, or a GPU context) through the
    fixture's steps; every dumped field bit for bit, every SIMPLE count."""
    fx = load(name)
    text = str(fx["text"])
    h = make(capi.parse_config3(text))
    iters = []
    dumps = dumped_steps(fx)
    np_ = 2 if f"s{dumps[0]}_pd1" in fx.files else 1
    for s in range(1, int(fx["steps"]) + 1):
        iters.append(h.step().simple_iterations)
        if s in dumps:
            for k, fid in FIELDS.items():
                if k.endswith("1") and np_ < 2:
                    continue
                ref = fx[f"s{s}_{k}"]
                got = h.get(fid)
                assert np.array_equal(got, ref), (name, s, k, float(np.abs(got - ref).max()))
    assert iters == list(fx["simple_iters"]), (name, iters, list(fx["simple_iters"]))
    return h


@pytest.mark.parametrize("name", CASES)
def test_oracle3d_matches_reference(name):
    h = check_case(name, pyoracle.Oracle3)
    h.close()


def digests(get):
    return {k: hashlib.sha256(np.ascontiguousarray(get(fid)).tobytes()).hexdigest()
            for k, fid in FIELDS.items() if k in ("u", "v", "w", "p", "flux", "pd0")}


def check_long(make, name="ref3d_cavity32_60"):
    fx = json.load(open(os.path.join(HERE, name + ".json")))
    h = make(capi.parse_config3(fx["text"]))
    iters = []
    for s in range(1, fx["steps"] + 1):
        iters.append(h.step().simple_iterations)
        if str(s) in fx["dumps"]:
            ref = fx["dumps"][str(s)]
            for k, d in digests(h.get).items():
                assert d == ref[k]["sha256"], (name, s, k)
    assert iters == fx["simple_iters"]
    return h


def test_oracle3d_long_cavity32():
    check_long(pyoracle.Oracle3).close()


def test_config3_reader():
    c = capi.parse_config3(workloads.dam3d(16, 4, 8) +
                           'set string condition_far "wall 0.5 -1 2"\n')
    assert (c.base.nx, c.base.ny, c.nz) == (16, 4, 8)
    assert (c.b_z, c.b1_z, c.gravity_z) == (1.0, 0.5, -9.81)
    assert list(c.bc_velocity[capi.HG_SIDE_FAR]) == [0.5, -1.0, 2.0]
    assert c.base.num_phases == 2 and c.base.density[1] == 1000.
    d = capi.parse_config3("")
    assert d.nz == 10 and d.b_z == 1.0 and list(d.bc_kind) == [0] * 6
