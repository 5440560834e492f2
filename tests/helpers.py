"""Shared test helpers: running the CLI in a run directory, the oracle's CLI path,
and the parity measures."""

import os
import sys
from contextlib import contextmanager

import numpy as np

from qpop_diffraction import SizeContext, AtomList, load_diffraction_hdf5
from qpop_diffraction.constants import OS0
from qpop_diffraction.h5lite import open_file, read_dataset
from qpop_diffraction.io_utils import (crop_slices, load_displacement_hdf5, load_material_field, load_phase_hdf5,
                                       load_platt_hdf5, load_region_hdf5)
from qpop_diffraction.json_input import load_input_json


@contextmanager
def cwd(path):
    old = os.getcwd()
    os.chdir(path)
    try:
        yield
    finally:
        os.chdir(old)


def run_cli(path, precision=32):
    """The package's qpop-diffraction in ``path``; returns (I (nz, ny, nx), [qx, qy, qz])."""
    from qpop_diffraction.cli import main
    with cwd(path):
        assert main(["--precision", str(precision)]) == 0
        with open_file("diffraction.h5", "r") as fh:
            I = read_dataset(fh, "intensity")
            q = [read_dataset(fh, f"q_{c}") for c in "xyz"]
    return I, q


def oracle_inputs(path):
    """The reference CLI's pre-processing (cli.py:149-217) with the reference-layout loaders.

    Returns (params, atoms, oPhase, oStruc, u, region, step) in the reference's (x, y, z) layout.
    """
    with cwd(path):
        cfg = load_input_json("input.json")
        lims = (cfg["x_limits"], cfg["y_limits"], cfg["z_limits"])
        params = SizeContext()
        fe_ids = params.read_json("system_setting.json")
        full = (params.nx, params.ny, params.nz)
        dz = params.lz / params.nz
        sl = crop_slices(*full, *lims)
        atoms = AtomList()
        assert atoms.read("parameter.atom.in", params.nPhase, params.nStruc)
        assert atoms.setup()
        oStruc = load_platt_hdf5(cfg["platt_file"], *full, *lims) / OS0
        with open_file(cfg["platt_file"], "r") as fh:
            step = int(fh.attrs["Timestep"])
        u = None
        if cfg["displacement_file"] is not None:
            u = load_displacement_hdf5(cfg["displacement_file"], *full, *lims)
        mat = load_material_field(os.path.join("data", "materialField_0.h5"), *full)
        oPhase = np.stack([mat[sl] == m for m in fe_ids]).astype(np.float64)
        region = oPhase.sum(axis=0).ravel()
        if cfg["phase_file"] is not None:
            oPhase = load_phase_hdf5(cfg["phase_file"], params.nPhase, *full, *lims)
        oPhase = oPhase.ravel()
        if cfg["region_file"] is not None:
            region *= load_region_hdf5(cfg["region_file"], *full, *lims).ravel()
        dc = params.reference_strain(np.diag(atoms.aC), params.temperature(step))[mat]
        uz_ref = ((np.cumsum(dc, axis=2) - 0.5 * dc) * dz)[sl].ravel()
        params.apply_slice(*lims)
        params.setup()
        if u is None:
            u = np.zeros(3 * params.n)
        u[2 * params.n:] += uz_ref
    return params, atoms, oPhase, oStruc, u, region, step


def oracle_run(path):
    """FP64 oracle intensity of a run directory, (nz, ny, nx) as the file stores it, and the q axes."""
    from qpd_oracle import diffraction_fp64
    params, atoms, oPhase, oStruc, u, region, _ = oracle_inputs(path)
    I, qx, qy, qz = diffraction_fp64(params, atoms, oPhase, oStruc, u, region)
    return I.transpose(2, 1, 0), [qx, qy, qz]


def golden_check(I, g, rtol_top, rtol_rel, atol_peak=1e-12):
    """Compare an intensity (nz, ny, nx) with a golden fixture dict.

    The stored float64 voxels (the 2048 largest and 2048 random ones) must agree
    to rtol_top with an absolute floor of atol_peak x the peak (the reference's own
    parity bound between its two implementations is rtol 1e-9, atol 1e-12 max,
    tests/test_mgpu.py:51-53); the whole float32 I/max map to rtol_rel of the peak.
    """
    flat = np.asarray(I, dtype=np.float64).ravel()
    atol = atol_peak * g["I_max"]
    np.testing.assert_allclose(flat[g["top_idx"]], g["top_val"], rtol=rtol_top, atol=atol)
    np.testing.assert_allclose(flat[g["samp_idx"]], g["samp_val"], rtol=rtol_top, atol=atol)
    rel = flat / g["I_max"]
    err = np.max(np.abs(rel - g["I_rel"].ravel().astype(np.float64)))
    assert err < rtol_rel, err


def dlog10(I, ref, floor_rel=1e-10):
    I = np.asarray(I, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    mask = ref >= floor_rel * ref.max()
    return float(np.max(np.abs(np.log10(np.maximum(I[mask], 1e-300)) - np.log10(ref[mask]))))


def rel_to_peak(I, ref):
    return float(np.max(np.abs(np.asarray(I) - np.asarray(ref))) / np.max(ref))


# The FP32 bounds: max |I - I_ref| <= 1e-5 of the peak, max |dlog10 I| <= 1e-3 over voxels
# >= 1e-10 of the peak, and the peak on the same voxel.
FP32_PEAK_TOL = 1e-5
FP32_DLOG_TOL = 1e-3


def check_fp32(I, ref, what=""):
    """Hold an FP32 intensity to the FP32 bounds against an FP64 one (same layout)."""
    rel, dl = rel_to_peak(I, ref), dlog10(I, ref)
    print(f"  fp32 {what}: max|dI|/peak {rel:.2e}, max|dlog10 I| {dl:.2e}", flush=True)
    assert rel <= FP32_PEAK_TOL, (what, rel)
    assert dl <= FP32_DLOG_TOL, (what, dl)
    assert np.argmax(I) == np.argmax(ref), what
