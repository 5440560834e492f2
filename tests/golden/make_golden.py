"""Write the golden fixtures of tests/golden/ from the reference's own NumPy path.

Run once, in the build container, with an interpreter that has h5py and can
import the reference package (the reference's package dir is linked into a
temporary sys.path entry; nothing is written under the reference tree):

    /opt/conda/bin/python3.9 tests/golden/make_golden.py [/path/to/reference]

For every case of golden_cases.py it writes the run directory, runs the
reference's ``qpop-diffraction`` main() in it with QPOP_DIFFRACTION_DEVICE=cpu,
and stores ``<case>.npz``:
  * I_rel      float32 (nz, ny, nx): /intensity divided by its maximum,
  * I_max      float64,
  * top_idx, top_val:   flat indices and float64 values of the 2048 largest voxels,
  * samp_idx, samp_val: 2048 seeded random voxels, float64,
  * q_x, q_y, q_z       float64 axes,
  * checksum   the input summary of synth.checksum.
It also copies the reference example 2_two_variant_twin (its inputs are fixture
data of the reference's tests/test_examples.py) and stores its intensity the
same way, and records whether the reference's own legacy comparison
(tests/test_examples.py:40-46) holds, in golden_meta.json.
"""

import json
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"

shim = tempfile.mkdtemp(prefix="qpd_refshim_")
os.symlink(os.path.join(REF, "src", "python"), os.path.join(shim, "qpop_diffraction"))
sys.path.insert(0, shim)
sys.path.insert(0, HERE)
os.environ["QPOP_DIFFRACTION_DEVICE"] = "cpu"

import h5py  # noqa: E402

from qpop_diffraction.cli import main as ref_main  # noqa: E402
import golden_cases  # noqa: E402


def run_reference(path):
    cwd = os.getcwd()
    argv = sys.argv
    try:
        os.chdir(path)
        sys.argv = ["qpop-diffraction"]
        rc = ref_main()
        assert rc == 0, rc
        with h5py.File(os.path.join(path, "diffraction.h5"), "r") as fh:
            return fh["intensity"][:], [fh["q_" + c][:] for c in "xyz"]
    finally:
        os.chdir(cwd)
        sys.argv = argv


def store(name, I, q, checksum):
    flat = I.ravel()
    top = np.argsort(flat)[::-1][:2048]
    rng = np.random.default_rng(12345)
    samp = rng.choice(flat.size, size=min(2048, flat.size), replace=False)
    np.savez_compressed(os.path.join(HERE, name + ".npz"),
                        I_rel=(I / I.max()).astype(np.float32), I_max=np.float64(I.max()),
                        top_idx=top.astype(np.int64), top_val=flat[top],
                        samp_idx=samp.astype(np.int64), samp_val=flat[samp],
                        q_x=q[0], q_y=q[1], q_z=q[2], checksum=np.asarray(checksum, dtype=np.float64))


def main():
    meta = {"reference": REF, "numpy": np.__version__, "cases": {}}
    work = tempfile.mkdtemp(prefix="qpd_golden_")
    for name, build in golden_cases.CASES.items():
        path = os.path.join(work, name)
        os.makedirs(path)
        run = build(h5py.File, path)
        I, q = run_reference(path)
        store(name, I, q, golden_cases.synth.checksum(run))
        meta["cases"][name] = {"shape_zyx": list(I.shape), "I_max": float(I.max())}
        print(name, I.shape, float(I.max()))

    # reference example with its own inputs
    ex = "2_two_variant_twin"
    dst = os.path.join(HERE, "examples", ex)
    if os.path.exists(dst):
        shutil.rmtree(dst)
    shutil.copytree(os.path.join(REF, "examples", ex), dst)
    for p in [os.path.join(dst, "diffraction.h5"), os.path.join(dst, "diffraction.xdmf")]:
        if os.path.exists(p):
            os.remove(p)
    tmp = os.path.join(work, ex)
    shutil.copytree(dst, tmp)
    I, q = run_reference(tmp)
    store("example_" + ex, I, q, [])
    meta["cases"]["example_" + ex] = {"shape_zyx": list(I.shape), "I_max": float(I.max())}

    # does the reference reproduce its own legacy fixture (tests/test_examples.py:40-46)?
    case = os.path.join("4_polar_vortex_PbTiO3", "disordered")
    tmp = os.path.join(work, "vortex")
    shutil.copytree(os.path.join(REF, "examples", case), tmp)
    I, q = run_reference(tmp)
    with h5py.File(os.path.join(REF, "examples", case, "reference_legacy.h5"), "r") as fh:
        legacy = fh["intensity"][:]
    strong = legacy > 1e-8 * legacy.max()
    ratio = I[strong] / legacy[strong]
    meta["legacy_vortex"] = {
        "passes_rtol_1e-4": bool(np.allclose(I[strong], legacy[strong], rtol=1e-4)),
        "ratio_min": float(ratio.min()), "ratio_max": float(ratio.max()),
        "peak_ref": float(I.max()), "peak_legacy": float(legacy.max()),
    }
    with open(os.path.join(HERE, "golden_meta.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    shutil.rmtree(work)
    shutil.rmtree(shim)
    print(json.dumps(meta["legacy_vortex"]))


if __name__ == "__main__":
    main()
