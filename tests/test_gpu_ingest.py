"""Field ingest on the GPU against the host preparation (cli.py:189-217).

For every golden run (films on substrates with a z crop, x/y crops, phase and region files,
the Voronoi two-phase run) and the reference example, the GPU-prepared P, U (with the u_z
reference ramp) and phase index equal the NumPy ones bit for bit, in FP32 and FP64, and the
command line writes the same intensity either way.
"""

import os
import sys

import numpy as np
import pytest

import golden_cases
from conftest import copy_example
from helpers import cwd, run_cli
from qpop_diffraction.h5lite import open_file

pytestmark = pytest.mark.gpu


def prepared(path, precision, device):
    from argparse import Namespace
    from qpop_diffraction.cli import prepare_run
    with cwd(path):
        args = Namespace(input="input.json", system="system_setting.json", atom="parameter.atom.in",
                         precision=precision, host_ingest=not device)
        params, atoms, f, step, out = prepare_run(args, device=device)
    get = lambda t: t.cpu().numpy() if hasattr(t, "cpu") else np.asarray(t)
    n = params.n
    return (get(f["P"]).reshape(3, n), get(f["U"]).reshape(3, n), get(f["phase"]).reshape(n))


def check_run(path):
    for precision in (32, 64):
        Ph, Uh, phh = prepared(path, precision, device=False)
        Pd, Ud, phd = prepared(path, precision, device=True)
        assert Pd.dtype == Ph.dtype
        assert np.array_equal(Pd, Ph), precision
        assert np.array_equal(Ud, Uh), (precision, np.max(np.abs(Ud - Uh)))
        assert np.array_equal(phd, phh), precision


@pytest.mark.parametrize("name", sorted(golden_cases.CASES))
def test_ingest_matches_host(tmp_path, name):
    print(f"ingest {name}", flush=True)
    golden_cases.CASES[name](open_file, str(tmp_path))
    check_run(tmp_path)


def test_ingest_example_and_cli_output(tmp_path):
    path = copy_example("2_two_variant_twin", tmp_path / "ex")
    check_run(path)
    I_dev, q_dev = run_cli(path, precision=32)
    from qpop_diffraction.cli import main
    with cwd(path):
        assert main(["--precision", "32", "--host-ingest"]) == 0
        with open_file("diffraction.h5", "r") as fh:
            from qpop_diffraction.h5lite import read_dataset
            I_host = read_dataset(fh, "intensity")
    assert np.array_equal(I_dev, I_host)
