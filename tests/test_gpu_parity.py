"""The HIP path (qpop-diffraction CLI and diffraction()) against the oracle and the golden vectors.

FP64 plans are held to the reference's own parity bound (rtol 1e-9, atol 1e-12 of
the peak, reference tests/test_mgpu.py:51-53).  FP32 plans are held to the FP32
bounds: max |I - I_ref| <= 1e-5 of the peak, max |dlog10 I| <= 1e-3 over voxels
>= 1e-10 of the peak, and the peak on the same voxel.
"""

import numpy as np
import pytest

import golden_cases
from conftest import copy_example, load_golden
from helpers import check_fp32, dlog10, golden_check, oracle_inputs, oracle_run, rel_to_peak, run_cli
from qpop_diffraction.h5lite import open_file

pytestmark = pytest.mark.gpu



@pytest.mark.parametrize("name", sorted(golden_cases.CASES))
def test_cli_fp64_matches_golden(tmp_path, name):
    print(f"cli fp64 {name}", flush=True)
    golden_cases.CASES[name](open_file, str(tmp_path))
    I, q = run_cli(tmp_path, precision=64)
    g = load_golden(name)
    golden_check(I, g, rtol_top=1e-9, rtol_rel=1e-7)
    for a, b in zip(q, (g["q_x"], g["q_y"], g["q_z"])):
        np.testing.assert_allclose(a, b, rtol=1e-14)


@pytest.mark.parametrize("name", sorted(golden_cases.CASES))
def test_cli_fp32_within_bulk_tolerance(tmp_path, name):
    print(f"cli fp32 {name}", flush=True)
    golden_cases.CASES[name](open_file, str(tmp_path))
    I, _ = run_cli(tmp_path, precision=32)
    ref, _ = oracle_run(tmp_path)
    check_fp32(I, ref)


@pytest.mark.parametrize("precision", [64, 32])
def test_example_two_variant_twin(tmp_path, precision):
    """Reference example 2_two_variant_twin (2000 x 20 x 1 grid, rotated aC, region file)."""
    print(f"example 2_two_variant_twin precision {precision}", flush=True)
    path = copy_example("2_two_variant_twin", tmp_path / "ex")
    I, _ = run_cli(path, precision=precision)
    g = load_golden("example_2_two_variant_twin")
    if precision == 64:
        golden_check(I, g, rtol_top=1e-9, rtol_rel=1e-7)
    else:
        ref, _ = oracle_run(path)
        check_fp32(I, ref)


REFERENCE_EXAMPLES = ["1_two_phase_mixture/anisotropic", "1_two_phase_mixture/isotropic",
                      "3_three_variant_hexagonal_twin", "4_polar_vortex_PbTiO3/ordered",
                      "4_polar_vortex_PbTiO3/disordered"]


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("name", REFERENCE_EXAMPLES)
def test_reference_examples(tmp_path, name, precision):
    """The reference's other examples (each of which its tests/test_examples.py:34-37 runs),
    from their own input files: 800 x 800 x 1 two-phase mixtures with phase-fraction and
    region files, a 1024 x 1024 x 1 three-variant hexagonal twin with displacements, and
    1 x 400 x 400 polar vortices.  The CLI against the oracle: FP64 at the reference's
    parity bound, FP32 at the bulk bound."""
    print(f"example {name} precision {precision}", flush=True)
    path = copy_example(name, tmp_path / "ex")
    I, q = run_cli(path, precision=precision)
    ref, qref = oracle_run(path)
    if precision == 64:
        np.testing.assert_allclose(I, ref, rtol=1e-9, atol=1e-12 * ref.max())
        for a, b in zip(q, qref):
            np.testing.assert_allclose(a, b, rtol=1e-14)
    else:
        check_fp32(I, ref)


@pytest.mark.parametrize("precision", [64, 32])
def test_dropin_diffraction_layout_matches_cli(tmp_path, precision):
    """diffraction() (z-fastest arrays) and the CLI (x-fastest files) give the same intensity."""
    from qpop_diffraction import diffraction
    print(f"drop-in diffraction precision {precision}", flush=True)
    golden_cases.case_twin_103(open_file, str(tmp_path))
    params, atoms, oPhase, oStruc, u, region, _ = oracle_inputs(tmp_path)
    I_xyz, qx, qy, qz = diffraction(params, atoms, oPhase, oStruc, u, region, precision=precision)
    I_cli, q = run_cli(tmp_path, precision=precision)
    scale = I_cli.max()
    np.testing.assert_allclose(I_xyz.transpose(2, 1, 0), I_cli, rtol=1e-5 if precision == 32 else 1e-10,
                               atol=1e-9 * scale if precision == 32 else 1e-13 * scale)
    for a, b in zip((qx, qy, qz), q):
        np.testing.assert_allclose(a, b, rtol=1e-15)


def test_dropin_fraction_weights(tmp_path):
    """Non-binary phase fractions take the fraction path of the source kernel."""
    from qpd_oracle import diffraction_fp64
    from qpop_diffraction import diffraction
    golden_cases.case_phase_region(open_file, str(tmp_path))
    params, atoms, oPhase, oStruc, u, region, _ = oracle_inputs(tmp_path)
    I, *_ = diffraction(params, atoms, oPhase, oStruc, u, region, precision=64)
    ref, *_ = diffraction_fp64(params, atoms, oPhase, oStruc, u, region)
    np.testing.assert_allclose(I, ref, rtol=1e-9, atol=1e-12 * ref.max())
