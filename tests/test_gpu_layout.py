"""Axis reversal (qpd_reverse_axes) and the CLI run z fastest.

The CLI lays FerroDyn fields out x fastest.  When no compile-time source kernel has the x
length and one has the z length, it reverses the fields to z fastest on the device, runs the
plan there, and reverses the intensity back before writing it.  The kernel is checked against
numpy's transpose, exactly, on ragged grids and every element size; the CLI with the layout
forced to z fastest is held to the golden vectors (FP64, the reference's bound) and to the
oracle (FP32, the bulk bound) on every golden case.
"""

import numpy as np
import pytest

import golden_cases
from conftest import load_golden
from helpers import check_fp32, golden_check, oracle_run, run_cli
from qpop_diffraction.h5lite import open_file

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,comp,dtype", [((37, 5, 70), 3, np.float32), ((1, 3, 33), 1, np.int8),
                                             ((64, 1, 1), 2, np.float64), ((33, 31, 65), 1, np.float64),
                                             ((100, 2, 7), 4, np.float32), ((32, 4, 32), 1, np.int8)])
def test_reverse_axes(dims, comp, dtype):
    import torch
    from qpop_diffraction._native import reverse_axes
    rng = np.random.default_rng(sum(dims))
    # C order (comp, d2, d1, d0): dims[0] fastest
    a = (rng.standard_normal((comp,) + dims[::-1]) * 100).astype(dtype)
    got = reverse_axes(torch.from_numpy(a).cuda(), dims).cpu().numpy().reshape((comp,) + tuple(dims))
    np.testing.assert_array_equal(got, a.transpose(0, 3, 2, 1))


def test_layout_choice(monkeypatch):
    from qpop_diffraction.cli import run_z_fastest
    monkeypatch.delenv("QPD_CLI_LAYOUT", raising=False)
    assert run_z_fastest(4000, 68)       # no 4000-point source kernel, a 68-point one
    assert not run_z_fastest(512, 104)   # both have one: x fastest as read
    assert not run_z_fastest(4000, 4001)  # neither
    monkeypatch.setenv("QPD_CLI_LAYOUT", "z")
    assert run_z_fastest(512, 104)
    monkeypatch.setenv("QPD_CLI_LAYOUT", "x")
    assert not run_z_fastest(4000, 68)


@pytest.mark.parametrize("name", sorted(golden_cases.CASES))
def test_cli_z_fastest(tmp_path, name, monkeypatch):
    print(f"cli z fastest {name}", flush=True)
    monkeypatch.setenv("QPD_CLI_LAYOUT", "z")
    golden_cases.CASES[name](open_file, str(tmp_path))
    I, _ = run_cli(tmp_path, precision=64)
    golden_check(I, load_golden(name), rtol_top=1e-9, rtol_rel=1e-7)
    I32, _ = run_cli(tmp_path, precision=32)
    ref, _ = oracle_run(tmp_path)
    check_fp32(I32, ref, name)
