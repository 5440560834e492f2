"""The hand-written Stockham 3-D FFT (qpd_fft3d) against numpy.fft, FP64 and FP32.

Lengths cover every butterfly radix the library implements (2, 3, 4, 5, 7, 8, 11, 13, 17),
the direct stages of larger prime factors (19, 23, 38 = 2 x 19, 46, 97, 4001, 5003), rows of
more than 4096 points (5003, 8000: sixteen points a thread), length-1 axes, and the axis
lengths of the workloads (104, 400, 512, 2000, 68).  FP64 must match to 1e-12 of the largest
output; FP32 (FP64-generated twiddles) to 2e-6 of it.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [  # (d2, d1, d0) = memory order, d0 fastest
    (1, 1, 1), (1, 1, 8), (4, 4, 4), (3, 5, 7), (11, 13, 17), (8, 16, 512), (13, 8, 104),
    (2, 17, 68), (1, 400, 400), (1, 20, 2000), (5, 9, 25), (6, 10, 12), (104, 4, 6), (17, 3, 11),
    (1, 1, 19), (3, 23, 38), (46, 2, 19), (97, 5, 23), (1, 2, 4001), (4001, 1, 3), (2, 4001, 2),
    (1, 2, 5003), (1, 2, 8000), (8000, 2, 1), (3, 8000, 2),
]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("precision", [64, 32])
def test_fft3d_matches_numpy(shape, precision):
    import torch
    from qpop_diffraction._native import fft3d
    print(f"fft3d shape={shape} precision={precision}", flush=True)
    rng = np.random.default_rng(hash(shape) % 2**32)
    nf = 2
    x = rng.standard_normal((nf,) + shape) + 1j * rng.standard_normal((nf,) + shape)
    dt = torch.complex128 if precision == 64 else torch.complex64
    t = torch.from_numpy(x).to(dtype=dt, device="cuda").contiguous()
    fft3d(t)
    torch.cuda.synchronize()
    got = t.cpu().numpy()
    ref = np.fft.fftn(x, axes=(1, 2, 3))
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    tol = 1e-12 if precision == 64 else 2e-6
    assert err < tol, err


def test_fft3d_rejects_unsupported_length():
    """Rows of more than 8192 points have no launch shape; the error names the length."""
    import torch
    from qpop_diffraction._native import QPDError, fft3d
    t = torch.zeros((1, 1, 8200), dtype=torch.complex64, device="cuda")
    with pytest.raises(QPDError, match="8200"):
        fft3d(t)
