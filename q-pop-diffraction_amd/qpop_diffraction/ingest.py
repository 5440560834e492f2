"""FerroDyn fields straight to the GPU (reference: src/python/cli.py:189-217, io_utils.py:30-92;
native counterpart src/heffte/qpop_diffraction_mgpu.cu:72-137 and kernels.cuh:34-50).

The reference converts every field to float64 on the host, transposes it, and builds the
one-hot phases and the u_z reference ramp with NumPy (a cumulative sum over the full stack).
Here the cropped components are read in the files' own precision (a float32 P file moves
4 bytes per value over PCIe, not 8), uploaded as they are, and one kernel
(``qpd_ingest_fields``) writes the plan inputs in file order: P / OS0, u / L0 with the ramp
on z, and the phase index of each cell.  Its arithmetic follows the host preparation step by
step, so the fields are bit-identical to those of ``cli.prepare_run(..., device=False)``.
"""

from __future__ import annotations

import ctypes

import numpy as np

from ._native import Ingest, check, current_stream, load_library
from .constants import L0, OS0
from .h5lite import open_file, read_dataset
from .io_utils import _checked, _prefix, crop_slices


def read_components(filepath, prefixes, nx, ny, nz, sl):
    """The three cropped components in file order and the files' own dtype."""
    with open_file(filepath, "r") as fh:
        prefix = _prefix(fh, prefixes)
        out = []
        for c in "xyz":
            name = f"{prefix}_{c}"
            _checked(fh, name, nx, ny, nz)
            out.append(np.ascontiguousarray(read_dataset(fh, name, (sl[2], sl[1], sl[0]), dtype=None)))
        return out


def read_material_columns(filepath, nx, ny, nz, sl):
    """materialField over the full z extent, cropped in x and y, as int32 (z, y, x).

    The reference reads it as float64 and casts to int (io_utils.py:89-92); ids are small
    integers, which both conversions keep exactly.
    """
    with open_file(filepath, "r") as fh:
        _checked(fh, "materialField", nx, ny, nz)
        raw = read_dataset(fh, "materialField", (slice(0, nz), sl[1], sl[0]), dtype=np.float64)
    return np.ascontiguousarray(raw.astype(np.int64).astype(np.int32))


def _float_bits(a):
    if a.dtype == np.float32:
        return 32
    if a.dtype == np.float64:
        return 64
    raise TypeError(f"field arrays must be float32 or float64, got {a.dtype}")


def ingest(P, U, material, dc, phase_of, z0, dz, precision=32, p_div=OS0, u_div=L0, stream=None):
    """Plan inputs (x fastest) from raw cropped components.

    P, U: three arrays each (z, y, x) cropped, float32 or float64 (U None: no displacement);
    material (nz_full, ny, nx) int32 over the full stack; dc (n_material,) float64;
    phase_of (n_material,) int8.  Arrays may be host numpy or device tensors.
    Returns device tensors P (3, n) real, U (3, n) float64, phase (n,) int8.
    """
    import torch
    dev = lambda a, dt=None: (a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a))).to(
        device="cuda", dtype=dt).contiguous()
    Pd = [dev(c) for c in P]
    nz, ny, nx = Pd[0].shape
    Ud = [dev(c) for c in U] if U is not None else None
    mat = dev(material, torch.int32)
    if mat.shape[1:] != (ny, nx):
        raise ValueError("material columns must match the crop in x and y")
    dcd = dev(np.asarray(dc, dtype=np.float64))
    phd = dev(np.asarray(phase_of, dtype=np.int8))
    n = nx * ny * nz
    real = torch.float32 if precision == 32 else torch.float64
    P_out = torch.empty((3, n), dtype=real, device="cuda")
    U_out = torch.empty((3, n), dtype=torch.float64, device="cuda")
    ph_out = torch.empty(n, dtype=torch.int8, device="cuda")
    bits = lambda t: {torch.float32: 32, torch.float64: 64}[t.dtype]
    arg = Ingest((ctypes.c_int64 * 3)(nx, ny, nz), int(mat.shape[0]), int(z0),
                 (ctypes.c_void_p * 3)(*[t.data_ptr() for t in Pd]), bits(Pd[0]), float(p_div),
                 (ctypes.c_void_p * 3)(*([t.data_ptr() for t in Ud] if Ud else [None, None, None])),
                 bits(Ud[0]) if Ud else 64, float(u_div), mat.data_ptr(), int(dcd.numel()), dcd.data_ptr(),
                 phd.data_ptr(), float(dz))
    check(load_library().qpd_ingest_fields(ctypes.byref(arg), int(precision), ctypes.c_void_p(P_out.data_ptr()),
                                           ctypes.c_void_p(U_out.data_ptr()), ctypes.c_void_p(ph_out.data_ptr()),
                                           stream if stream is not None else current_stream()))
    return P_out, U_out, ph_out
