"""The CPU oracle against what pins it to the reference (CPU only).

* Golden fixtures dumped by the reference module itself (built from its
  sources by oracle/ref.mk; tests/golden/make_fixtures.py): the oracle must
  reproduce every field bit for bit, with the same SIMPLE iteration counts
  and the same number of SOR/Jacobi sweeps per pressure solve.
* The reference's shipped sample run (examples/cavity/sample/exp.log): the
  1000 per-iteration convergence indicators Rs, printed with 6 significant
  digits, must be reproduced to that precision, and the final fields written
  through the library's .vts writer must reproduce the sample's field file
  (exp.field.0.vts) byte for byte.
"""
import os

import numpy as np
import pytest

from hydro_amd import capi
from oracle import pyoracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["ref_cavity16", "ref_cavity64", "ref_dam64x32", "ref_cavity32_jacobi",
         "ref_dam32x16_dtauto", "ref_cavity32_heat", "ref_dam32x16_heat",
         "ref_cavity32_simpler", "ref_dam32x16_simpler", "ref_cavity32_simpler_pinned",
         "ref_channel32x16", "ref_channel32x16_jacobi", "ref_cavity32_lu_relaxed",
         "ref_cavity32_lu_relaxed_tight", "ref_cavity32_vel_lu_relaxed",
         "ref_dam32x16_vel_jacobi"]

# Solvers that print "iter = , diff = " after each solve (linear.hpp:647,712,779)
PRINTING = {capi.SOLVERS[n] for n in ("gauss_seidel", "jacobi", "lu_relaxed")}

FIELDS = [("u", capi.HG_FIELD_VELOCITY_X), ("v", capi.HG_FIELD_VELOCITY_Y),
          ("p", capi.HG_FIELD_PRESSURE), ("flux", capi.HG_FIELD_VOLUME_FLUX)]


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_reference_fixture(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    text = str(z["text"])
    steps = int(z["steps"])
    o = pyoracle.Oracle(capi.parse_config(text))
    dumps = sorted({int(k[1:k.index("_")]) for k in z.files
                    if k.startswith("s") and k[1:2].isdigit()})
    ref_iters = list(z["simple_iters"])
    ref_sweeps = list(z["pressure_sweeps"])
    # per SIMPLE iteration the velocity solves (two components) print before
    # the pressure solve(s); SIMPLER solves the pressure twice
    vp = 2 if o.cfg.linear_solver_velocity in PRINTING else 0
    npr = 2 if o.cfg.simpler else 1
    pos = 0
    for s in range(1, steps + 1):
        st = o.step()
        assert st.simple_iterations == ref_iters[s - 1], (name, s)
        n = st.simple_iterations
        chunk = ref_sweeps[pos:pos + n * (vp + npr)]
        pressure = [chunk[q * (vp + npr) + vp + r] for q in range(n) for r in range(npr)]
        assert st.pressure_solve_iters_total == sum(pressure), (name, s)
        pos += n * (vp + npr)
        if o.cfg.heat_enable and o.cfg.linear_solver_heat != capi.SOLVERS["lu"]:
            pos += 1  # the heat solve's own sweep count follows the step's
        if s in dumps:
            for fname, f in FIELDS:
                assert np.array_equal(o.get(f), z[f"s{s}_{fname}"]), (name, s, fname)
            for i in range(o.cfg.num_phases):
                assert np.array_equal(o.get(capi.HG_FIELD_VOLUME_FRACTION_0 + i),
                                      z[f"s{s}_alpha{i}"]), (name, s, i)
                assert np.array_equal(o.get(capi.HG_FIELD_PARTIAL_DENSITY_0 + i),
                                      z[f"s{s}_pd{i}"]), (name, s, i)
            if o.cfg.heat_enable:
                assert np.array_equal(o.get(capi.HG_FIELD_TEMPERATURE), z[f"s{s}_T"]), (name, s)
    assert pos == len(ref_sweeps)


@pytest.mark.parametrize("name", ["ref_cavity32_vel_lu_relaxed", "ref_dam32x16_vel_jacobi"])
def test_oracle_velocity_sweeps_match_reference(name):
    """The velocity solves' own sweep counts (the two prints before each
    pressure solve) over the first step's SIMPLE iterations."""
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    o = pyoracle.Oracle(capi.parse_config(str(z["text"])))
    sweeps = list(z["pressure_sweeps"])
    o.start_step()
    q = 0
    while not o.is_converged():
        st = o.iterate()
        assert [st.velocity_solve_iters[0], st.velocity_solve_iters[1],
                st.pressure_solve_iters] == sweeps[q:q + 3], (name, q)
        q += 3
    assert q == 3 * int(z["simple_iters"][0])


def test_oracle_reproduces_cavity_sample_log(tmp_path):
    text = open(os.path.join(GOLDEN, "cavity_sample.hydroconf")).read()
    rs_ref = np.loadtxt(os.path.join(GOLDEN, "cavity_sample_rs.txt"))
    o = pyoracle.Oracle(capi.parse_config(text))
    o.start_step()
    out = []
    while not o.is_converged():
        out.append(o.iterate().convergence_indicator)
    assert len(out) == len(rs_ref) == 1000
    rel = np.abs(np.array(out) - rs_ref) / rs_ref
    # 6 significant digits in the log: rounding error <= 5e-6 relative
    assert rel.max() < 5e-6, rel.max()
    # ... and as printed, every one of them
    assert ["%g" % v for v in out] == ["%g" % v for v in rs_ref]
    o.finish_step()
    check_sample_frame(o, tmp_path)


def check_sample_frame(o, tmp_path):
    """The sample's field frame (examples/cavity/sample/exp.field.0.vts,
    written by the reference at the end of the run) byte for byte from the
    oracle's fields through hg_vts_write: node x, y; cell velocity, pressure,
    volume fraction (the sample's output_* flags)."""
    import gzip
    want = gzip.open(os.path.join(GOLDEN, "sample_exp.field.0.vts.gz"), "rt").read()
    n = o.cfg.nx
    xn = np.array([o.cfg.a[0] + (i / n) * (o.cfg.b[0] - o.cfg.a[0]) for i in range(n + 1)])
    yn = np.array([o.cfg.a[1] + (j / n) * (o.cfg.b[1] - o.cfg.a[1]) for j in range(n + 1)])
    gx, gy = np.meshgrid(xn, yn)
    path = str(tmp_path / "exp.field.0.vts")
    capi.vts_write(path, xn, yn, [("x", gx.ravel()), ("y", gy.ravel())],
                   [("velocity_x", o.get(capi.HG_FIELD_VELOCITY_X)),
                    ("velocity_y", o.get(capi.HG_FIELD_VELOCITY_Y)),
                    ("pressure", o.get(capi.HG_FIELD_PRESSURE)),
                    ("volume_fraction_0", o.get(capi.HG_FIELD_VOLUME_FRACTION_0))])
    assert open(path).read() == want


def field_digests(get, num_phases):
    """SHA-256 of each FP64 field's bytes, keyed like the reference dump."""
    import hashlib
    out = {}
    arrays = [(name, get(f)) for name, f in FIELDS]
    for i in range(num_phases):
        arrays.append((f"alpha{i}", get(capi.HG_FIELD_VOLUME_FRACTION_0 + i)))
        arrays.append((f"pd{i}", get(capi.HG_FIELD_PARTIAL_DENSITY_0 + i)))
    for name, a in arrays:
        out[name] = (hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes())
                     .hexdigest(), float(np.linalg.norm(a)))
    return out


def check_long_fixture(fix, digests_at_step, step_stats):
    """Compares a run with a long-trajectory fixture: SIMPLE iterations per
    step, total pressure sweeps, and every dumped field bit for bit (digest),
    reporting the relative L2-norm difference of any mismatch."""
    assert [s.simple_iterations for s in step_stats] == fix["simple_iters"]
    assert sum(s.pressure_solve_iters_total for s in step_stats) == fix["pressure_sweeps_total"]
    for step, fields in fix["dumps"].items():
        got = digests_at_step[int(step)]
        bad = {k: abs(got[k][1] - v["l2"]) / max(v["l2"], 1e-300)
               for k, v in fields.items() if k in got and got[k][0] != v["sha256"]}
        assert not bad, (step, bad)


@pytest.mark.parametrize("name", ["ref_cavity64_200", "ref_dam256x128_50", pytest.param(
    "ref_dam256x128_500", marks=pytest.mark.skipif(
        os.environ.get("HYDRO_LONG_C2") != "1",
        reason="C2 500 steps (about 40 minutes on the CPU): set HYDRO_LONG_C2=1"))])
def test_oracle_long_trajectory(name):
    import json
    fix = json.load(open(os.path.join(GOLDEN, name + ".json")))
    o = pyoracle.Oracle(capi.parse_config(fix["text"]))
    stats, digests = [], {}
    for s in range(1, fix["steps"] + 1):
        stats.append(o.step())
        if str(s) in fix["dumps"]:
            digests[s] = field_digests(o.get, o.cfg.num_phases)
    check_long_fixture(fix, digests, stats)
