"""Host-side workload generators, the byte manifest and the option defaults
(CPU only)."""
import re

import numpy as np
import pytest

from hydro_amd import capi, manifest, workloads
from hydro_amd.hydro import batch_run, train_opts


def dims(text):
    return int(re.search(r"Nx (\d+)", text).group(1)), int(re.search(r"Ny (\d+)", text).group(1))


def test_c5_cases_are_seeded_and_in_range():
    a = workloads.c5_cases(count=64, seed=2026)
    b = workloads.c5_cases(count=64, seed=2026)
    assert a == b and len(a) == 64
    assert a != workloads.c5_cases(count=64, seed=1)
    cells = 0
    for t in a:
        nx, ny = dims(t)
        assert nx in range(64, 513, 32)
        assert ny in (nx, nx // 2)
        cells += nx * ny
        cfg = capi.parse_config(t)          # every case reads as reference .hydroconf text
        assert cfg.unsupported_flags == 0
        assert cfg.nx == nx and cfg.ny == ny
    assert 3e6 <= cells <= 7e6            # the batch size of the workload definition


def test_rt_volume_fraction_is_exact_for_a_flat_interface():
    a = workloads.rt_volume_fraction(16, 32, amp=0.0, noise=0.0).reshape(32, 16)
    assert np.array_equal(a[:16], np.zeros((16, 16)))
    assert np.array_equal(a[16:], np.ones((16, 16)))


def test_rt_volume_fraction_area_matches_the_integral():
    # mean of alpha = area above the interface / domain area; the cosine
    # integrates to zero over whole periods
    a = workloads.rt_volume_fraction(64, 128, amp=0.01, noise=0.0)
    assert abs(a.mean() - 0.5) < 1e-12


def test_manifest_cg_iteration_is_112_bytes_per_cell():
    nx, ny = 1024, 2048
    b = manifest.cg_iteration_bytes(nx, ny) / (nx * ny)
    assert 112.0 <= b < 112.1


def test_train_opts_defaults_mirror_the_library():
    o = train_opts()
    assert (o.rhs_count, o.mg_iters, o.steps, o.seed) == (4, 12, 300, 1)
    assert abs(o.lr - 1e-3) < 1e-9
    assert o.flags == capi.HG_TRAIN_TIED


def test_batch_run_rejects_bad_arguments_without_a_device():
    with pytest.raises(Exception):
        batch_run([], -1)
