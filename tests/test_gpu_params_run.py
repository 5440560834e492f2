"""A prepared run (``--mgpu`` then ``params_run``) against the plain command line, as the
reference's test_mgpu.py checks its multi-GPU program against the Python path.

On the planted runs (bulk, film on substrate with the u_z ramp, cropped film with the ramp
summed over the full stack, phase and region files) and the reference example, both routes
write the same q axes and the same /intensity: bit for bit in FP32 and FP64 here (the
reference's bound between its two programs is rtol 1e-9, atol 1e-12 of the peak).
"""

import numpy as np
import pytest

import golden_cases
from conftest import copy_example
from helpers import cwd
from qpop_diffraction.h5lite import open_file, read_dataset

pytestmark = pytest.mark.gpu


def outputs():
    with open_file("diffraction.h5", "r") as fh:
        return read_dataset(fh, "intensity"), [read_dataset(fh, f"q_{c}") for c in "xyz"]


def compare(path, precision):
    from qpop_diffraction.cli import main
    from qpop_diffraction.params_run import main as params_main
    with cwd(path):
        assert main(["--precision", str(precision)]) == 0
        I_ref, q_ref = outputs()
        assert main(["--mgpu"]) == 0
        with open_file("diffraction.h5", "r") as fh:
            assert "intensity" not in fh
        assert params_main(["diffraction_params.h5", "--precision", str(precision)]) == 0
        I, q = outputs()
    for a, b in zip(q, q_ref):
        np.testing.assert_array_equal(a, b)
    assert I.shape == I_ref.shape
    np.testing.assert_allclose(I, I_ref, rtol=1e-9, atol=1e-12 * I_ref.max())
    assert np.array_equal(I, I_ref), np.max(np.abs(I - I_ref)) / I_ref.max()


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("name", sorted(golden_cases.CASES))
def test_params_run_matches_cli(tmp_path, name, precision):
    print(f"params_run {name} FP{precision}", flush=True)
    golden_cases.CASES[name](open_file, str(tmp_path))
    compare(tmp_path, precision)


def test_params_run_example(tmp_path):
    path = copy_example("2_two_variant_twin", tmp_path / "ex")
    compare(path, 32)
