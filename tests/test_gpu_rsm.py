"""Reciprocal-space maps from device projections against the oracle's restatement of
rsm.py:66-113 (every voxel histogrammed), for default bins, a bin override and the mean
aggregation, in both grid layouts; and the qpop-diffraction-rsm command on a written file."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bins,aggregation", [(None, "sum"), (None, "mean"), (7, "sum"), (9, "mean")])
def test_device_maps_match_oracle(bins, aggregation):
    import torch
    from qpd_oracle import rsm_map
    from qpop_diffraction.rsm import PAIR_DEFINITIONS, reciprocal_space_maps
    rng = np.random.default_rng(1)
    nx, ny, nz = 40, 36, 22
    Ixyz = rng.random((nx, ny, nz)) ** 4 * 1e6
    q = [np.fft.fftshift(np.fft.fftfreq(n, 0.39)) * 2 * np.pi + off for n, off in ((nx, 0.0), (ny, 0.0), (nz, 32.2))]
    x_fast = torch.from_numpy(np.ascontiguousarray(Ixyz.transpose(2, 1, 0))).cuda()  # memory (nx, ny, nz) dims
    z_fast = torch.from_numpy(np.ascontiguousarray(Ixyz)).cuda()                      # memory (nz, ny, nx) dims
    for I, dims, axis in ((x_fast, (nx, ny, nz), (0, 1, 2)), (z_fast, (nz, ny, nx), (2, 1, 0))):
        maps = reciprocal_space_maps(I, dims, axis, q, bins, aggregation)
        for a, b, stem in PAIR_DEFINITIONS:
            ca, cb, ref = rsm_map(Ixyz, q, a, b, bins, aggregation)
            ga, gb, got = maps[stem]
            np.testing.assert_allclose(ga, ca, rtol=0, atol=0)
            np.testing.assert_allclose(gb, cb, rtol=0, atol=0)
            np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * ref.max())


def test_rsm_command(tmp_path):
    import torch
    from qpd_oracle import rsm_map
    from qpop_diffraction.io_utils import write_diffraction_output
    from qpop_diffraction.rsm import main
    rng = np.random.default_rng(2)
    I = rng.random((10, 8, 6))  # (nx, ny, nz)
    q = [np.linspace(-1, 1, 10), np.linspace(-1, 1, 8), np.linspace(31, 33, 6)]
    write_diffraction_output(str(tmp_path / "d"), I, *q, 0)
    assert main(["-d", str(tmp_path / "d.h5"), "-o", str(tmp_path / "maps")]) == 0
    with np.load(tmp_path / "maps" / "qx_qz.npz", allow_pickle=False) as z:
        ref = rsm_map(I, q, 0, 2)[2]
        np.testing.assert_allclose(z["intensity"], ref, rtol=1e-12)
        np.testing.assert_allclose(z["log10_intensity"], np.log10(ref), rtol=1e-12)
        assert str(z["q_units"]) == "nm^-1"
