"""The output sessions on the GPU path against files the reference wrote.

* tests/golden/ref_output_*.npz: every file the reference module's writers
  produced (ParaView frames and collection, scalar table, iteration history)
  when run as the reference console runs it: step, advance the time,
  write_results (TModule::cycle, module.cpp:46-73).  The same loop over
  hg_step and hg_output_write_results must write the same files: frames,
  collection and history byte for byte (the exact-order path reproduces
  the reference's fields bit for bit), the scalar table to its printed
  precision (the GPU's statistics are reductions in another order).
* The reference's cavity sample run (examples/cavity/sample): one step of
  1000 SIMPLE iterations; its field frame and iteration history byte for
  byte.
"""
import gzip
import os

import numpy as np
import pytest

from hydro_amd import capi
from hydro_amd.hydro import Hydro, Unsupported
from tests.test_output import OUTPUT_FIXTURES, load_output_fixture, scalar_table

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def run_cycle(h, out, t_end, max_steps):
    """hydro::step, the time advance and write_results(finished), as the
    reference's module loop; stops once the time reaches T."""
    for _ in range(max_steps):
        h.step()
        finished = h.progress().module_time >= t_end
        out.write_results(finished)
        if finished:
            break


def tables_agree(want, got):
    """Scalar tables: same header and rows; numbers equal as printed, or
    within the last printed digit (sums in a different order)."""
    names_w, rows_w = scalar_table(want)
    names_g, rows_g = scalar_table(got)
    assert names_w == names_g and len(rows_w) == len(rows_g)
    cols = np.array([[float(v) for v in r] for r in rows_w])
    scale = np.abs(cols).max(axis=0)
    for rw, rg in zip(rows_w, rows_g):
        for k, (a, b) in enumerate(zip(rw, rg)):
            if a == b:
                continue
            fa, fb = float(a), float(b)
            assert abs(fa - fb) <= 1e-5 * max(abs(fa), abs(fb)) or \
                abs(fa - fb) <= 1e-12 * scale[k], (names_w[k], a, b)


@pytest.mark.parametrize("name", OUTPUT_FIXTURES)
def test_output_files_match_reference_writers(gpu, tmp_path, monkeypatch, name):
    text, out_text, steps, files = load_output_fixture(name)
    print(f"case: {name}, {steps} steps", flush=True)
    monkeypatch.chdir(tmp_path)
    with Hydro(text) as h:
        out = h.open_output(out_text)
        run_cycle(h, out, out.oc.T, steps + 5)
        out.close()
        assert h.progress().steps == steps
    written = sorted(os.listdir(tmp_path))
    assert written == sorted(files)
    for f, want in files.items():
        got = open(tmp_path / f).read()
        if f.endswith(".dat"):
            tables_agree(want, got)
        else:
            assert got == want, f


def test_cavity_sample_frame_and_history(gpu, tmp_path, monkeypatch):
    """The reference's sample run: T = dt, so one step then a forced write;
    max_frame_index 0 (no frame at the start), the history of step 0."""
    print("case: cavity sample 64x64, 1 step of 1000 SIMPLE iterations", flush=True)
    text = open(os.path.join(GOLDEN, "cavity_sample.hydroconf")).read()
    monkeypatch.chdir(tmp_path)
    with Hydro(text) as h:
        out = h.open_output(text)
        assert out.oc.iter_history_enable == 1 and out.oc.max_frame_index == 0
        run_cycle(h, out, out.oc.T, 2)
        out.close()
        assert h.progress().iteration == 1000
    for f in ("exp.field.0.vts", "exp.iter_history.plt"):
        want = gzip.open(os.path.join(GOLDEN, "sample_" + f + ".gz"), "rt").read()
        assert open(tmp_path / f).read() == want, f
    assert sorted(os.listdir(tmp_path)) == ["exp.field.0.vts", "exp.field.pvd",
                                            "exp.iter_history.plt", "exp.scalar.dat"]


def test_running_stats_match_oracle(gpu):
    """hg_calc_stats' boundary integrals and mesh offset against the
    oracle's (which prints as the reference's table) on the moving-mesh
    cavity, every step."""
    from oracle import pyoracle
    text, _, _, _ = load_output_fixture("ref_output_cavity16_meshvel")
    print("case: cavity 16x16 meshvel, 12 steps, stats vs oracle", flush=True)
    o = pyoracle.Oracle(capi.parse_config(text))
    with Hydro(text) as h:
        for step in range(12):
            g, r = h.stats(), o.stats()
            for f in ("volume", "mass", "cx", "cy", "vx", "vy", "volume_in", "volume_out",
                      "mass_in", "mass_out", "pd_min", "pd_max"):
                a, b = getattr(g, f)[0], getattr(r, f)[0]
                assert abs(a - b) <= 1e-12 * max(abs(b), 1e-3), (step, f, a, b)
            pr = h.progress()
            assert pr.steps == step
            h.step()
            o.step()


def test_output_refuses_entries_it_cannot_supply(gpu, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    with Hydro("set int Nx 16\nset int Ny 16\n") as h:
        for key in ("output_volume_source", "output_mass_source_0"):
            with pytest.raises(Unsupported, match=key):
                h.open_output(f"set bool {key} 1\n")
        out = h.open_output("set bool no_output 1\n")
        h.step()
        out.write_results(True)
        out.close()
    assert os.listdir(tmp_path) == []
