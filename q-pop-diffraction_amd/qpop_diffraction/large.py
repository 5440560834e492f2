"""FP64 intensity of a crop around G for volumes too large for an FP64 plan.

An FP64 plan of a 4000 x 4000 x 68 volume needs 16 complex128 work fields per cell
(about 280 GB), so the FP64 comparison of such a volume is computed for the central
crop only, in the layout with z fastest (memory dims (nz, ny, nx), axes (2, 1, 0)):

  * passes 0 (z) and 1 (y) act at fixed x, so they run exactly on x slabs of the volume:
    a slab plan has the full z and y lengths and tables; its pass-0 coefficient carries
    1/n of the slab, which is rescaled to 1/n of the volume;
  * of each slab's pass-1 output (two fields per z-class, the input of the last pass)
    only the crop's ky frequencies are kept, stored at ky mod Ly for an Ly-point crop;
  * pass 2 (x, the full length) then runs on a reduced grid (nz, Ly, nx), whose fftshift
    along y places frequency ky at crop index ky + Ly/2, exactly as the crop of the full
    fftshifted axis; the x crop is taken from its output.

The result equals the crop of the FP64 volume (the same kernels, tables and order of
operations as an FP64 plan of the whole grid).  It is the FP64 reference that the FP32
result of a large volume is compared with.
"""

from __future__ import annotations

import numpy as np

from ._native import Plan


def _plan(dims, axis, params, tables, precision):
    # the passes in memory order (z, y, x): the slabs are cut along the last pass's axis
    R, f, su, G, qC, V = tables
    return Plan(dims, axis, (params.dx, params.dy, params.dz), R, f, su, G, qC, V, precision, order="memory")


def crop_indices(n, m):
    """Indices of the central m of an fftshifted axis of n points, and their FFT-order frequencies."""
    start = n // 2 - m // 2
    s = np.arange(start, start + m)
    k = (s - n // 2)  # frequency index relative to zero, in [-m/2, m/2)
    return s, k


def fp64_crop(vol, params, tables, crop=(512, 512), slab=200, log=None):
    """FP64 intensity on the central crop[0] (x) x crop[1] (y) x all z of a z-fastest volume.

    vol: object with dims (nz, ny, nx), axis (2, 1, 0), device tensors P (3, n), U (3, n)
    FP64 and phase (n,) int8.  Returns a device float64 tensor shaped (cx, cy, nz) in the
    fftshifted order of the full volume's crop.
    """
    import torch
    nz, ny, nx = vol.dims
    if tuple(vol.axis) != (2, 1, 0):
        raise ValueError("fp64_crop expects the z-fastest layout (axes (2, 1, 0))")
    cx, cy = min(crop[0], nx), min(crop[1], ny)
    if cy % 2 or cx % 2:
        raise ValueError("crop sizes must be even")
    sy, ky = crop_indices(ny, cy)
    sx, _ = crop_indices(nx, cx)
    ky_full = torch.as_tensor(np.mod(ky, ny), device="cuda")
    ky_red = torch.as_tensor(np.mod(ky, cy), device="cuda")
    n_full = nz * ny * nx
    # reduced grid for the last pass: its workspace holds the pass-1 outputs
    red = _plan((nz, cy, nx), (2, 1, 0), params, tables, 64)
    n_red = nz * cy * nx
    work_red = torch.zeros(red.workspace_bytes // 16, dtype=torch.complex128, device="cuda")
    # workspace of an FP64 plan: the 2 nZ pass-1 outputs (pass 2's inputs) first, then the
    # pass-0 outputs
    z0 = torch.zeros((3, 1), dtype=torch.float64, device="cuda")
    i0 = torch.zeros(1, dtype=torch.int8, device="cuda")
    n_in2 = int(round((red.pass_bytes_per_cell(2, z0, z0, i0) - 8.0) / 16.0))
    plane = nz * ny
    for x0 in range(0, nx, slab):
        x1 = min(nx, x0 + slab)
        w = x1 - x0
        sp = _plan((nz, ny, w), (2, 1, 0), params, tables, 64)
        n_s = nz * ny * w
        P = vol.P[:, x0 * plane:x1 * plane].to(torch.float64).contiguous()
        U = vol.U[:, x0 * plane:x1 * plane].contiguous()
        ph = vol.phase[x0 * plane:x1 * plane].contiguous()
        work = torch.empty(sp.workspace_bytes // 16, dtype=torch.complex128, device="cuda")
        dummy = torch.empty(1, dtype=torch.float64, device="cuda")
        sp.run_pass(0, P, U, ph, dummy, work)
        sp.run_pass(1, P, U, ph, dummy, work)
        out1 = work[:n_in2 * n_s].view(n_in2, w, ny, nz)
        sel = out1.index_select(2, ky_full) * (n_s / n_full)  # (fields, w, cy, nz)
        dst = work_red[:n_in2 * n_red].view(n_in2, nx, cy, nz)
        dst[:, x0:x1].index_copy_(2, ky_red, sel)
        del work, P, U, ph, out1, sel
        sp.close()
        if log:
            log(f"fp64 crop: slab {x1}/{nx}")
    I = torch.empty(n_red, dtype=torch.float64, device="cuda")
    red.run_pass(2, z0, z0, i0, I, work_red)
    torch.cuda.synchronize()
    red.close()
    Ired = I.view(nx, cy, nz)
    return Ired[torch.as_tensor(sx, device="cuda")].contiguous()


def crop_of(I, dims, crop=(512, 512)):
    """The same crop of a full z-fastest intensity (device tensor of n values): (cx, cy, nz)."""
    import torch
    nz, ny, nx = dims
    cx, cy = min(crop[0], nx), min(crop[1], ny)
    sy, _ = crop_indices(ny, cy)
    sx, _ = crop_indices(nx, cx)
    v = I.view(nx, ny, nz)
    return v[torch.as_tensor(sx, device="cuda")][:, torch.as_tensor(sy, device="cuda")].contiguous()
