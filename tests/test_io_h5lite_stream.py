"""The streamed output writer and the HDF5 slab writes it relies on (CPU)."""

import numpy as np
import pytest

from qpop_diffraction.h5lite import open_file, read_dataset
from qpop_diffraction.io_utils import write_diffraction_output, write_diffraction_streamed


@pytest.mark.parametrize("force", ["", "1"])
def test_streamed_writer_matches_plain_writer(tmp_path, monkeypatch, force):
    import torch
    if force:
        monkeypatch.setenv("QPOP_FORCE_H5LITE", "1")
    rng = np.random.default_rng(0)
    nz, ny, nx = 7, 5, 6
    I = rng.random((nz, ny, nx))
    q = [np.linspace(-1, 1, n) for n in (nx, ny, nz)]
    write_diffraction_output(str(tmp_path / "a"), I, *q, 3, zyx=True)
    # slabs of two planes, so the last slab is partial
    write_diffraction_streamed(str(tmp_path / "b"), torch.from_numpy(I.copy()), *q, 3, slab_bytes=2 * ny * nx * 8)
    with open_file(str(tmp_path / "a.h5"), "r") as fa, open_file(str(tmp_path / "b.h5"), "r") as fb:
        for name in ("intensity", "q_x", "q_y", "q_z"):
            assert np.array_equal(read_dataset(fa, name), read_dataset(fb, name)), name
        assert int(fa.attrs["Timestep"]) == int(fb.attrs["Timestep"]) == 3
    assert (tmp_path / "b.xdmf").read_text() == (tmp_path / "a.xdmf").read_text().replace("a.h5", "b.h5")
