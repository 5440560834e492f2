"""FerroDyn field input and diffraction output (reference: src/python/io_utils.py:16-141,186-207).

FerroDyn stores each field component as a dataset shaped (nz, ny, nx), x fastest,
in SI units.  Two families of loaders are offered:

* the reference's own, which return the cropped field in (x, y, z) C order,
  float64, flattened (``load_platt_hdf5`` and friends) - the input layout of
  ``diffraction()``;
* file-order loaders (``read_field``, ``read_vector``) that keep (z, y, x) order
  and the dtype asked for.  The GPU path consumes that layout directly, so the
  command-line run never transposes a field on the host.

Output: ``{base}.h5`` with ``/intensity`` shaped (nz, ny, nx) float64, the axes
``/q_x``, ``/q_y``, ``/q_z`` (nm^-1), attributes ``Timestep`` and ``q_units``,
and the XDMF 3DRectMesh sidecar ``{base}.xdmf``.
"""

from __future__ import annotations

import os

import numpy as np

from .constants import L0
from .h5lite import open_file, read_dataset


def crop_slices(nx, ny, nz, x_lim, y_lim, z_lim):
    """Inclusive 0-based limits (None = whole axis) as (x, y, z) slices."""
    out = []
    for lim, n in ((x_lim, nx), (y_lim, ny), (z_lim, nz)):
        out.append(slice(0, n) if lim is None else slice(lim[0], lim[1] + 1))
    return tuple(out)


def _checked(fh, name, nx, ny, nz):
    shape = tuple(fh[name].shape)
    if shape != (nz, ny, nx):
        raise ValueError(
            f"Dataset '{name}' in {fh.filename} has shape {shape}, expected "
            f"(nz, ny, nx) = ({nz}, {ny}, {nx}) from system_setting.json.  Files "
            f"written before 2026-05-26 are (nx, ny, nz): convert them with "
            f"qpop_monitor.migrate_hdf5_layout.")


def read_field(fh, name, nx, ny, nz, sl, dtype=np.float64):
    """One component, cropped, in file order (z, y, x), converted to dtype."""
    _checked(fh, name, nx, ny, nz)
    return np.ascontiguousarray(read_dataset(fh, name, (sl[2], sl[1], sl[0]), dtype=dtype))


def _prefix(fh, prefixes):
    for p in prefixes:
        if f"{p}_x" in fh:
            return p
    raise StopIteration(f"none of {prefixes} found in {fh.filename}")


def read_vector(filepath, prefixes, nx, ny, nz, x_lim=None, y_lim=None, z_lim=None, dtype=np.float64):
    """``{prefix}_x/y/z`` cropped, as (3, nz_s, ny_s, nx_s) in file order."""
    sl = crop_slices(nx, ny, nz, x_lim, y_lim, z_lim)
    with open_file(filepath, "r") as fh:
        prefix = _prefix(fh, prefixes)
        return np.stack([read_field(fh, f"{prefix}_{c}", nx, ny, nz, sl, dtype) for c in "xyz"])


def _xyz(field_zyx):
    return field_zyx.transpose(2, 1, 0)


def load_platt_hdf5(filepath, nx, ny, nz, x_lim=None, y_lim=None, z_lim=None):
    """P in C/m^2, cropped, (x, y, z) C order flattened as [Px | Py | Pz] (io_utils.py:54-60)."""
    v = read_vector(filepath, ("P_Lattice", "PLatt"), nx, ny, nz, x_lim, y_lim, z_lim)
    return np.stack([_xyz(c) for c in v]).ravel()


def load_displacement_hdf5(filepath, nx, ny, nz, x_lim=None, y_lim=None, z_lim=None):
    """u in nm (files hold metres), (x, y, z) C order flattened as [ux | uy | uz] (io_utils.py:63-69)."""
    v = read_vector(filepath, ("Displacement",), nx, ny, nz, x_lim, y_lim, z_lim)
    return np.stack([_xyz(c) for c in v]).ravel() / L0


def load_phase_hdf5(filepath, nPhase, nx, ny, nz, x_lim=None, y_lim=None, z_lim=None):
    """Phase fractions phase_1 .. phase_<nPhase>, cropped, as (nPhase, nx_s, ny_s, nz_s) (io_utils.py:72-78)."""
    sl = crop_slices(nx, ny, nz, x_lim, y_lim, z_lim)
    with open_file(filepath, "r") as fh:
        return np.stack([_xyz(read_field(fh, f"phase_{p + 1}", nx, ny, nz, sl)) for p in range(nPhase)])


def load_region_hdf5(filepath, nx, ny, nz, x_lim=None, y_lim=None, z_lim=None):
    """Scattering weight ``region``, cropped, as (nx_s, ny_s, nz_s) (io_utils.py:81-86)."""
    sl = crop_slices(nx, ny, nz, x_lim, y_lim, z_lim)
    with open_file(filepath, "r") as fh:
        return _xyz(read_field(fh, "region", nx, ny, nz, sl))


def read_material_field(filepath, nx, ny, nz):
    """materialField, uncropped, in file order (z, y, x), as int64."""
    full = crop_slices(nx, ny, nz, None, None, None)
    with open_file(filepath, "r") as fh:
        return read_field(fh, "materialField", nx, ny, nz, full, np.float64).astype(np.int64)


def load_material_field(filepath, nx, ny, nz):
    """materialField uncropped, integer, (x, y, z) order (io_utils.py:89-92)."""
    return _xyz(read_material_field(filepath, nx, ny, nz))


# -- output -------------------------------------------------------------------

def write_diffraction_axes(base_name, qx, qy, qz, step):
    """``{base}.h5`` without /intensity (attributes and q axes) plus the XDMF sidecar (io_utils.py:108-141)."""
    nx, ny, nz = len(qx), len(qy), len(qz)
    print(f"Writing diffraction output → {base_name}.h5")
    with open_file(f"{base_name}.h5", "w") as fh:
        fh.attrs["Timestep"] = step
        fh.attrs["q_units"] = "nm^-1"
        fh["q_x"] = np.asarray(qx, dtype=np.float64)
        fh["q_y"] = np.asarray(qy, dtype=np.float64)
        fh["q_z"] = np.asarray(qz, dtype=np.float64)

    h5 = os.path.basename(f"{base_name}.h5")

    def item(dims, path):
        return f'<DataItem Dimensions="{dims}" NumberType="Float" Precision="8" Format="HDF">{h5}:/{path}</DataItem>'

    xdmf = "\n".join([
        '<?xml version="1.0" ?>',
        '<Xdmf Version="3.0">',
        '  <Domain>',
        '    <Grid Name="DiffractionPattern" GridType="Uniform">',
        f'      <Topology TopologyType="3DRectMesh" Dimensions="{nz} {ny} {nx}"/>',
        '      <Geometry GeometryType="VXVYVZ">',
        f'        {item(nx, "q_x")}',
        f'        {item(ny, "q_y")}',
        f'        {item(nz, "q_z")}',
        '      </Geometry>',
        '      <Attribute Name="Intensity" AttributeType="Scalar" Center="Node">',
        f'        {item(f"{nz} {ny} {nx}", "intensity")}',
        '      </Attribute>',
        '    </Grid>',
        '  </Domain>',
        '</Xdmf>',
        '',
    ])
    with open(f"{base_name}.xdmf", "w", encoding="utf-8") as fh:
        fh.write(xdmf)
    print(f"Writing ParaView sidecar  → {base_name}.xdmf")


def write_diffraction_output(base_name, I, qx, qy, qz, step, zyx=False):
    """Write ``{base}.h5`` and ``{base}.xdmf`` (io_utils.py:95-105).

    I is (nx, ny, nz) as ``diffraction()`` returns it, or already (nz, ny, nx)
    when ``zyx`` is true; /intensity is stored (nz, ny, nx).
    """
    write_diffraction_axes(base_name, qx, qy, qz, step)
    data = I if zyx else np.ascontiguousarray(I.transpose(2, 1, 0))
    with open_file(f"{base_name}.h5", "a") as fh:
        fh["intensity"] = np.asarray(data, dtype=np.float64)


def write_mgpu_params(path, full, limits, spacing, G, qC, V, R, fAtom, su, phase_of, dc,
                      platt_file, displacement_file, material_file, phase_file, region_file, output_file):
    """The inputs of a prepared run, ``{output_file}_params.h5`` (io_utils.py:144-183).

    Datasets grid, crop (x0, x1, y0, y1, z0, z1 inclusive), spacing (nm), G, qC (nm^-1), V
    (nm^3), R (nAtom, 3), fAtom_re / fAtom_im (nPhase, nAtom), su (3, nStruc, nPhase, nAtom),
    phase_of and dc per material, u_scale (nm per m); file paths as string attributes.  The
    reference's multi-GPU program reads this file (params.hpp:13-108); ``params_run`` runs it
    on one GPU.
    """
    with open_file(str(path), "w") as fh:
        fh["grid"] = np.asarray(full, dtype=np.int64)
        fh["crop"] = np.asarray(limits, dtype=np.int64)
        fh["spacing"] = np.asarray(spacing, dtype=np.float64)
        fh["G"] = np.asarray(G, dtype=np.float64)
        fh["qC"] = np.asarray(qC, dtype=np.float64)
        fh["V"] = np.array([V], dtype=np.float64)
        fh["R"] = np.asarray(R, dtype=np.float64)
        fh["fAtom_re"] = np.ascontiguousarray(np.real(fAtom), dtype=np.float64)
        fh["fAtom_im"] = np.ascontiguousarray(np.imag(fAtom), dtype=np.float64)
        fh["su"] = np.asarray(su, dtype=np.float64)
        fh["phase_of"] = np.asarray(phase_of, dtype=np.int32)
        fh["dc"] = np.asarray(dc, dtype=np.float64)
        fh["u_scale"] = np.array([1.0 / L0])
        with open_file(str(platt_file), "r") as fp:
            platt_prefix = next(p for p in ("P_Lattice", "PLatt") if f"{p}_x" in fp)
        for key, value in (("platt_file", platt_file), ("platt_prefix", platt_prefix),
                           ("displacement_file", displacement_file), ("material_file", material_file),
                           ("phase_file", phase_file), ("region_file", region_file), ("output_file", output_file)):
            if value is not None:
                fh.attrs[key] = np.bytes_(str(value))
    print(f"Writing prepared-run parameters → {path}")


def read_mgpu_params(path) -> dict:
    """``{output}_params.h5`` as written by ``write_mgpu_params`` (params.hpp:98-140); missing file
    attributes are None."""
    out = {}
    with open_file(str(path), "r") as fh:
        for key in ("grid", "crop", "spacing", "G", "qC", "V", "R", "fAtom_re", "fAtom_im", "su", "phase_of", "dc",
                    "u_scale"):
            if key not in fh:
                raise KeyError(f"Missing dataset {key} in the prepared-run parameters {path}")
            dt = np.int64 if key in ("grid", "crop", "phase_of") else np.float64
            out[key] = read_dataset(fh, key, dtype=dt)
        for key in ("platt_file", "platt_prefix", "displacement_file", "material_file", "phase_file", "region_file",
                    "output_file"):
            v = fh.attrs[key] if key in fh.attrs else None
            out[key] = (v.decode() if isinstance(v, (bytes, np.bytes_)) else (None if v is None else str(v)))
    if out["su"].ndim != 4 or out["su"].shape[1] != 3:
        raise ValueError("the prepared run expects nStruc = 3 (P_x, P_y, P_z)")
    return out


_STAGING = {}


def _staging(n, pinned):
    """Two reusable host buffers of at least n float64 values (page-locked for device copies)."""
    import torch
    key = bool(pinned)
    have = _STAGING.get(key)
    if have is None or have[0].numel() < n:
        have = [torch.empty(n, dtype=torch.float64, pin_memory=key) for _ in range(2)]
        _STAGING[key] = have
    return have


def write_diffraction_streamed(base_name, I, qx, qy, qz, step, slab_bytes=64 << 20, axes=True):
    """``write_diffraction_output`` for an intensity still on the GPU, written as it arrives.

    I: device tensor of nz ny nx float64 values in (z, y, x) order (the command line's
    layout).  The file is the one ``write_diffraction_output`` writes; the intensity goes in
    slabs of whole z planes: the copy of slab k into a pinned host buffer runs on its own
    stream while slab k - 1 is written from the other buffer.  Returns the seconds spent.
    """
    import time
    import torch
    t0 = time.perf_counter()
    nz, ny, nx = (len(qz), len(qy), len(qx))
    flat = I.reshape(nz, ny * nx)
    if axes:
        write_diffraction_axes(base_name, qx, qy, qz, step)
    planes = max(1, min(nz, slab_bytes // max(1, ny * nx * 8)))
    on_gpu = flat.is_cuda
    bufs = _staging(planes * ny * nx, on_gpu)
    bufs = [b[:planes * ny * nx].view(planes, ny * nx) for b in bufs]
    copy = torch.cuda.Stream() if on_gpu else None
    events = [torch.cuda.Event() if on_gpu else None for _ in range(2)]
    if on_gpu:
        copy.wait_stream(torch.cuda.current_stream())
    with open_file(f"{base_name}.h5", "a") as fh:
        if "intensity" in fh:
            del fh["intensity"]
        ds = fh.create_dataset("intensity", shape=(nz, ny, nx), dtype=np.float64)
        slabs = [(z0, min(nz, z0 + planes)) for z0 in range(0, nz, planes)]

        def fetch(k):
            z0, z1 = slabs[k]
            b = bufs[k % 2]
            if on_gpu:
                with torch.cuda.stream(copy):
                    b[:z1 - z0].copy_(flat[z0:z1], non_blocking=True)
                    events[k % 2].record(copy)
            else:
                b[:z1 - z0].copy_(flat[z0:z1])

        def store(k):
            z0, z1 = slabs[k]
            if on_gpu:
                events[k % 2].synchronize()
            ds[z0:z1] = bufs[k % 2][:z1 - z0].numpy().reshape(z1 - z0, ny, nx)

        fetch(0)
        for k in range(1, len(slabs)):
            fetch(k)
            store(k - 1)
        store(len(slabs) - 1)
    return time.perf_counter() - t0


def load_diffraction_hdf5(filepath) -> dict:
    """Read a diffraction.h5 back: intensity as (1, nx, ny, nz) and the q grid (io_utils.py:186-207)."""
    with open_file(filepath, "r") as fh:
        I = read_dataset(fh, "intensity", dtype=np.float64).transpose(2, 1, 0)
        q = [read_dataset(fh, f"q_{c}", dtype=np.float64) for c in "xyz"]
        units = fh.attrs["q_units"]
    if isinstance(units, bytes):
        units = units.decode()
    shape = I.shape
    grids = [np.broadcast_to(qd.reshape([-1 if d == i else 1 for i in range(3)]), shape) for d, qd in enumerate(q)]
    return {
        "dims": shape,
        "intensity": I[np.newaxis],
        "log10_intensity": np.log10(I)[np.newaxis],
        "q_x": q[0], "q_y": q[1], "q_z": q[2],
        "q_vector": np.stack(grids),
        "q_units": str(units),
    }
