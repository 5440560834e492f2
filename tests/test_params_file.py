"""``qpop-diffraction --mgpu``: the prepared-run file and the output skeleton (cli.py:105-147,
io_utils.py:144-183), on the CPU.

For each planted run of the reference's test_mgpu.py (bulk, film on substrate, cropped film,
phase and region files) the file holds the full grid, the inclusive crop, the cropped spacing,
G and qC from the oracle's reciprocal_centre, the crystal tables, phase_of and dc per material;
the output skeleton holds the oracle's q axes and no intensity.  A write/read round trip keeps
every value and the absent file attributes.
"""

import json
import os

import numpy as np
import pytest

import golden_cases
from helpers import cwd
from qpop_diffraction import AtomList, SizeContext
from qpop_diffraction.cli import main
from qpop_diffraction.constants import L0, OS0
from qpop_diffraction.h5lite import open_file, read_dataset
from qpop_diffraction.io_utils import read_mgpu_params, write_mgpu_params


@pytest.mark.parametrize("name", sorted(golden_cases.CASES))
def test_prepared_run_file(tmp_path, name, capsys):
    from qpd_oracle import q_axes, reciprocal_centre
    golden_cases.CASES[name](open_file, str(tmp_path))
    with cwd(tmp_path):
        assert main(["--mgpu"]) == 0
        cfg = json.load(open("input.json"))
        prm = read_mgpu_params("diffraction_params.h5")
        params = SizeContext()
        fe_ids = params.read_json("system_setting.json")
        atoms = AtomList()
        assert atoms.read("parameter.atom.in", params.nPhase, params.nStruc) and atoms.setup()
    assert "params_run" in capsys.readouterr().out
    full = (params.nx, params.ny, params.nz)
    assert tuple(prm["grid"]) == full
    lims = [cfg.get(k) for k in ("x_limits", "y_limits", "z_limits")]
    crop = sum(([0, n - 1] if l is None else list(l) for l, n in zip(lims, full)), [])
    assert list(prm["crop"]) == crop
    spacing = np.array([params.lx / params.nx, params.ly / params.ny, params.lz / params.nz])
    ncrop = np.array(crop[1::2]) - np.array(crop[0::2]) + 1
    np.testing.assert_allclose(prm["spacing"], spacing, rtol=1e-15)
    # G, qC of the cropped box (the strain average does not depend on the box)
    G, qC = reciprocal_centre(atoms.aC, atoms.q00, params.strainAvg)
    np.testing.assert_allclose(prm["G"], G, rtol=1e-14)
    np.testing.assert_allclose(prm["qC"], qC, rtol=1e-14)
    assert prm["V"][0] == pytest.approx(np.linalg.det(atoms.aC), rel=1e-15)
    np.testing.assert_array_equal(prm["R"], atoms.xAtom @ atoms.aC)
    np.testing.assert_array_equal(prm["fAtom_re"] + 1j * prm["fAtom_im"], atoms.fAtom)
    np.testing.assert_array_equal(prm["su"], np.stack([atoms.su1, atoms.su2, atoms.su3]) / OS0)
    phase_of = np.full(len(params.names), -1)
    phase_of[fe_ids] = np.arange(len(fe_ids))
    np.testing.assert_array_equal(prm["phase_of"], phase_of)
    with open_file(prm["platt_file"], "r") as fh:
        step = int(fh.attrs["Timestep"])
    np.testing.assert_array_equal(prm["dc"], params.reference_strain(np.diag(atoms.aC), params.temperature(step)))
    assert prm["u_scale"][0] == 1.0 / L0
    assert prm["output_file"] == str((tmp_path / "diffraction.h5").resolve())
    assert prm["platt_prefix"] in ("P_Lattice", "PLatt")
    for key in ("phase_file", "region_file", "displacement_file"):
        assert (prm[key] is None) == (cfg.get(key) in (None, "")), key
    with open_file(prm["output_file"], "r") as fh:
        assert "intensity" not in fh
        q = [read_dataset(fh, f"q_{c}") for c in "xyz"]
        assert int(fh.attrs["Timestep"]) == step
    dx = prm["spacing"]
    for a, b in zip(q, q_axes(tuple(ncrop), tuple(dx), qC)):
        np.testing.assert_allclose(a, b, rtol=1e-14)


def test_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    platt = tmp_path / "p.h5"
    with open_file(str(platt), "w") as fh:
        fh["PLatt_x"] = np.zeros((1, 1, 1))
    args = dict(full=(40, 32, 24), limits=(2, 37, 0, 31, 4, 17), spacing=(0.4, 0.41, 0.39),
                G=rng.standard_normal(3), qC=rng.standard_normal(3), V=0.0625, R=rng.standard_normal((5, 3)),
                fAtom=rng.standard_normal((2, 5)) + 1j * rng.standard_normal((2, 5)),
                su=rng.standard_normal((3, 3, 2, 5)), phase_of=np.array([-1, 0, 1]), dc=rng.standard_normal(3),
                platt_file=platt, displacement_file=None, material_file=tmp_path / "m.h5", phase_file=None,
                region_file=tmp_path / "r.h5", output_file=tmp_path / "o.h5")
    write_mgpu_params(tmp_path / "x_params.h5", **args)
    got = read_mgpu_params(tmp_path / "x_params.h5")
    assert tuple(got["grid"]) == args["full"] and tuple(got["crop"]) == args["limits"]
    for key in ("spacing", "G", "qC", "R", "su", "phase_of", "dc"):
        np.testing.assert_array_equal(got[key], np.asarray(args[key]), err_msg=key)
    assert got["V"][0] == args["V"]
    np.testing.assert_array_equal(got["fAtom_re"] + 1j * got["fAtom_im"], args["fAtom"])
    assert got["platt_prefix"] == "PLatt"
    assert got["displacement_file"] is None and got["phase_file"] is None
    assert got["region_file"] == str(args["region_file"]) and got["output_file"] == str(args["output_file"])


def test_missing_dataset(tmp_path):
    with open_file(str(tmp_path / "bad.h5"), "w") as fh:
        fh["grid"] = np.array([1, 1, 1])
    with pytest.raises(KeyError, match="crop"):
        read_mgpu_params(tmp_path / "bad.h5")
