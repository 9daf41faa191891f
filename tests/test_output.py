"""The output row on the CPU: the output keys, the .vts writer, and the
oracle's running statistics against what the reference's own writers wrote.

* Fixtures tests/golden/ref_output_*.npz hold every file the reference
  module's writers produced for two cases (tests/golden/make_fixtures.py):
  the oracle's statistics at each scalar row must print as the reference's
  .dat table does (the running boundary in/out integrals and the mesh
  offset meshpos included: the cavity case moves the mesh).
* The .vts writer (hg_vts_write, host code of libhydrogpu.so) fed with the
  oracle's fields of the reference's cavity sample run must reproduce the
  sample's field file byte for byte (tests/test_oracle_reference.py).
"""
import os

import numpy as np
import pytest

from hydro_amd import capi
from oracle import pyoracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
OUTPUT_FIXTURES = ["ref_output_dam32x16", "ref_output_cavity16_meshvel",
                   "ref_output_dam32x16_dtauto", "ref_output_dam32x16_heat",
                   "ref_output_channel32x16"]


def load_output_fixture(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    files = {str(f): str(c) for f, c in zip(z["files"], z["contents"])}
    return str(z["text"]), str(z["output"]), int(z["steps"]), files


def scalar_table(text):
    lines = text.splitlines()
    names = lines[0].split()
    rows = [l.split() for l in lines[1:] if l.strip()]
    assert all(len(r) == len(names) for r in rows)
    return names, rows


def test_output_config_defaults_are_general_hydroconf():
    oc = capi.default_output_config()
    assert oc.abi_version == capi.HG_ABI_VERSION
    assert oc.name == b"exp" and oc.field_output_format == b"paraview"
    assert oc.plt_title == b"Template"
    assert (oc.max_frame_index, oc.max_frame_scalar_index) == (100, 10000)
    assert tuple(oc.output_factor) == (1, 1) and oc.T == 1.0
    assert (oc.no_output, oc.no_mesh_output, oc.iter_history_enable) == (0, 0, 0)
    on = {capi.OUTPUT_ENTRIES[e] for e in range(len(capi.OUTPUT_ENTRIES)) if oc.output[e]}
    # examples/general.hydroconf:156-188
    assert on == {"x", "y", "velocity_x", "velocity_y", "pressure", "volume_fraction_0",
                  "volume_fraction_1"}


def test_output_config_parse():
    oc = capi.parse_output_config("""
set string name run1
set string plt_title "Some title"
set double T 2.5
set int max_frame_index 7
set int output_factor_y 3
set bool output_divergence 1
set bool output_x 0
set bool output_partial_density_2 1
set string field_output_format tecplot_binary
set int Nx 64
""")
    assert oc.name == b"run1" and oc.plt_title == b"Some title"
    assert oc.T == 2.5 and oc.max_frame_index == 7 and tuple(oc.output_factor) == (1, 3)
    assert oc.field_output_format == b"tecplot_binary"
    ent = capi.OUTPUT_ENTRIES
    assert oc.output[ent.index("divergence")] == 1 and oc.output[ent.index("x")] == 0
    assert oc.output[ent.index("partial_density_2")] == 1
    with pytest.raises(ValueError):
        capi.parse_output_config("set double T x\n")


def test_vts_writer_layout(tmp_path):
    """Structure of a frame: extents, PointData / CellData / Points, values
    as %g with a space after each, the point list ending without a line
    break (output_paraview.hpp:45-131)."""
    xn, yn = np.array([0., 0.5, 1.]), np.array([0., 1. / 3.])
    path = str(tmp_path / "f.vts")
    capi.vts_write(path, xn, yn, [("x", [0., 0.5, 1., 0., 0.5, 1.])],
                   [("p", [1e-7, -2.5])])
    want = ('<?xml version="1.0"?>\n'
            '<VTKFile type="StructuredGrid" version="0.1" byte_order="LittleEndian">\n'
            '  <StructuredGrid WholeExtent="0 2 0 1 0 0">\n'
            '    <Piece Extent="0 2 0 1 0 0">\n'
            '      <PointData>\n'
            '        <DataArray Name="x" NumberOfComponents="1" type="Float32" format="ascii">\n'
            '0 0.5 1 0 0.5 1 \n'
            '        </DataArray>\n'
            '      </PointData>\n'
            '      <CellData>\n'
            '        <DataArray Name="p" NumberOfComponents="1" type="Float32" format="ascii">\n'
            '1e-07 -2.5 \n'
            '        </DataArray>\n'
            '      </CellData>\n'
            '      <Points>\n'
            '        <DataArray Name="mesh" NumberOfComponents="3" type="Float32" format="ascii">\n'
            '0 0 0 0.5 0 0 1 0 0 0 0.333333 0 0.5 0.333333 0 1 0.333333 0 '
            '        </DataArray>\n'
            '      </Points>\n'
            '    </Piece>\n'
            '  </StructuredGrid>\n'
            '</VTKFile>\n')
    assert open(path).read() == want


def test_vts_writer_long_arrays_keep_order(tmp_path):
    """Arrays longer than one formatting chunk (4096 values) are formatted on
    several threads; the text must equal the one-thread formatting."""
    nx, ny = 150, 70
    rng = np.random.default_rng(3)
    v = rng.standard_normal(nx * ny) * 10.0 ** rng.integers(-12, 12, nx * ny)
    path = str(tmp_path / "g.vts")
    capi.vts_write(path, np.linspace(0, 1, nx + 1), np.linspace(0, 2, ny + 1), [], [("v", v)])
    lines = open(path).read().splitlines()
    k = lines.index('        <DataArray Name="v" NumberOfComponents="1" type="Float32" '
                    'format="ascii">')
    assert lines[k + 1] == "".join("%g " % a for a in v)


def test_vts_writer_rejects_bad_arguments(tmp_path):
    lib = capi.load()
    assert lib.hg_vts_write(None, 1, 1, None, None, 0, None, None, 0, None, None) == \
        capi.HG_ERR_BAD_ARG
    with pytest.raises(RuntimeError, match="HG_ERR_IO"):
        capi.vts_write(str(tmp_path / "missing_dir" / "f.vts"), [0., 1.], [0., 1.], [], [])


@pytest.mark.parametrize("name", OUTPUT_FIXTURES)
def test_oracle_stats_print_as_reference_scalar_table(name):
    """Every row of the reference's .dat table against the oracle's state at
    that time, printed the same way: time, "s", the convergence indicator,
    and CalcStat's per-phase values (hydro2d.hpp:825-893).  The time column
    is the module's clock (module.cpp:74-81): t = the previous t + the dt in
    force before the step, which trails the fluid time under dt_auto."""
    text, _, steps, files = load_output_fixture(name)
    dat = [v for f, v in files.items() if f.endswith(".dat")]
    names, rows = scalar_table(dat[0])
    o = pyoracle.Oracle(capi.parse_config(text))
    np_ = o.cfg.num_phases

    def row(t, s, ind):
        st = o.stats()
        vals = [t, s, ind]
        for i in range(np_):
            vals += [st.mass[i], st.mass_in[i], st.mass_out[i]]
        for i in range(np_):
            vals += [st.volume[i], st.volume_in[i], st.volume_out[i]]
        for i in range(np_):
            vals += [st.pd_min[i], st.pd_max[i]]
        for i in range(np_):
            vals += [st.cx[i], st.cy[i], st.vx[i], st.vy[i]]
        return ["%g" % v for v in vals]

    assert len(names) == 3 + np_ * 12
    got = {rows[0][0]: row(0.0, 0, 1.0)}
    t, tnp = 0.0, 0.0 + o.cfg.dt
    for _ in range(steps):
        st = o.step()
        t = tnp
        tnp = t + st.dt
        got["%g" % t] = row(t, st.simple_iterations, st.convergence_indicator)
    for r in rows:
        assert r[0] in got, r[0]
        assert got[r[0]] == r, (r[0], [(n, a, b) for n, a, b in zip(names, got[r[0]], r)
                                       if a != b])
    # the moving-mesh case has nonzero boundary fluxes and a shifted centre
    if name.endswith("meshvel"):
        last = dict(zip(names, rows[-1]))
        assert float(last["volume_in_0"]) > 0 and float(last["volume_out_0"]) > 0
