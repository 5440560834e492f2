#!/usr/bin/env python3
"""``qpop-diffraction``: the drop-in command line (reference: src/python/cli.py:66-94,149-225).

Reads input.json, system_setting.json (with materials/ and data/materialField_0.h5
beside it) and parameter.atom.in, and writes {output_file}.h5 (/intensity shaped
(nz, ny, nx), /q_x, /q_y, /q_z) and {output_file}.xdmf, as the reference does.

The fields stay in FerroDyn's file order (x fastest) from the HDF5 reader to the
GPU and back to the writer: the plan is built for memory dims (nx, ny, nz) with
physical axes (0, 1, 2), so the intensity comes out already shaped (nz, ny, nx).
``--mgpu`` writes the reference's prepared-run file ({output_file}_params.h5) and the
output skeleton; ``python -m qpop_diffraction.params_run`` then computes it on one
MI355X, which holds the grids the reference's multi-GPU program was written for.
"""

import argparse
import os
import sys
import time
from pathlib import Path

import numpy as np

from .atom_list import AtomList
from .constants import L0, OS0
from ._native import fast_lengths, reverse_axes
from .diffraction import Engine, Volume, default_precision, q_axes, reciprocal_centre
from .h5lite import open_file
from .io_utils import (crop_slices, read_field, read_material_field, read_vector, write_diffraction_axes,
                       write_diffraction_streamed, write_mgpu_params)
from .json_input import load_input_json
from .size_context import SizeContext


def _parse_args(argv=None) -> argparse.Namespace:
    parser = argparse.ArgumentParser(
        description="Q-POP Diffraction Simulation for Q-POP-FerroDyn output (MI355X)",
        formatter_class=argparse.ArgumentDefaultsHelpFormatter,
    )
    parser.add_argument("--input", "-i", default="input.json", metavar="JSON",
                        help="Path to input.json (P_Lattice/Displacement paths, slice limits, output name).")
    parser.add_argument("--system", "-s", default="system_setting.json", metavar="JSON",
                        help="Path to the FerroDyn system_setting.json of the run.")
    parser.add_argument("--atom", "-a", default="parameter.atom.in", metavar="FILE",
                        help="Path to parameter.atom.in.")
    parser.add_argument("--precision", type=int, choices=(32, 64), default=None,
                        help="Working precision: 32 (complex FP32, FP64 phase reduction) or 64. "
                             "Default: QPOP_DIFFRACTION_PRECISION or 32.")
    parser.add_argument("--mgpu", action="store_true",
                        help="Only prepare the run: write {output_file}_params.h5 and the {output_file}.h5 skeleton "
                             "(as the reference does for its multi-GPU program); run it on one GPU with "
                             "`python -m qpop_diffraction.params_run {output_file}_params.h5`.")
    parser.add_argument("--host-ingest", action="store_true",
                        help="Prepare P, u, the phases and the u_z ramp with NumPy on the host instead of on the GPU "
                             "(the results are bit-identical).")
    return parser.parse_args(argv)


def prepare_run(args, device=None):
    """Everything before the GPU call: config, crystal, fields in file order (cli.py:149-217).

    device: prepare P, u, the phase index and the u_z ramp on the GPU (``ingest``) from the
    raw cropped components; otherwise with NumPy on the host.  Default: the GPU unless
    args.host_ingest.  Returns (params, atoms, fields dict, step, output base name) or None.
    """
    if device is None:
        device = not getattr(args, "host_ingest", False)
    cfg = load_input_json(args.input)
    lims = (cfg["x_limits"], cfg["y_limits"], cfg["z_limits"])

    params = SizeContext()
    fe_ids = params.read_json(args.system)
    full = (params.nx, params.ny, params.nz)
    dz = params.lz / params.nz
    sl = crop_slices(*full, *lims)

    atoms = AtomList()
    if not atoms.read(args.atom, params.nPhase, params.nStruc) or not atoms.setup():
        return None

    if getattr(args, "mgpu", False):
        return _prepare_params(args, cfg, params, fe_ids, atoms, lims)
    if device:
        return _prepare_on_device(args, cfg, params, fe_ids, atoms, full, dz, sl, lims)
    real = np.float32 if args.precision == 32 else np.float64
    print(f"Loading P_Lattice file: {cfg['platt_file']}")
    P = read_vector(cfg["platt_file"], ("P_Lattice", "PLatt"), *full, *lims, dtype=real)
    if OS0 != 1.0:
        P = P / OS0
    with open_file(cfg["platt_file"], "r") as fh:
        step = int(fh.attrs["Timestep"])

    if cfg["displacement_file"] is not None:
        print(f"Loading Displacement file: {cfg['displacement_file']}")
        U = read_vector(cfg["displacement_file"], ("Displacement",), *full, *lims, dtype=np.float64) / L0
    else:
        print("No displacement_file provided — using zero displacement field.")
        U = np.zeros((3,) + P.shape[1:], dtype=np.float64)

    # phases: cells of non-FE materials do not scatter (cli.py:189-200)
    mat_file = Path(args.system).parent / "data" / "materialField_0.h5"
    print(f"Loading material field: {mat_file}")
    mat = read_material_field(mat_file, *full)  # (nz, ny, nx)
    zyx = (sl[2], sl[1], sl[0])
    phase_of = np.full(len(params.names), -1, dtype=np.int8)
    phase_of[fe_ids] = np.arange(len(fe_ids), dtype=np.int8)
    phase = phase_of[mat[zyx]]
    fraction = None
    if cfg["phase_file"] is not None:
        print(f"Loading phase fractions: {cfg['phase_file']}")
        with open_file(cfg["phase_file"], "r") as fh:
            fraction = np.stack([read_field(fh, f"phase_{p + 1}", *full, sl, real) for p in range(params.nPhase)])
    region = None
    if cfg["region_file"] is not None:
        print(f"Loading scattering weight: {cfg['region_file']}")
        with open_file(cfg["region_file"], "r") as fh:
            region = read_field(fh, "region", *full, sl, real)

    # u_z reference ramp over the full stack, then cropped (cli.py:202-217)
    print(f"\nTimestep {step}:")
    dc = params.reference_strain(np.diag(atoms.aC), params.temperature(step))[mat]
    uz_ref = ((np.cumsum(dc, axis=0) - 0.5 * dc) * dz)[zyx]
    U[2] += uz_ref

    print("\nApplying domain slice:")
    params.apply_slice(*lims)
    params.setup()
    fields = {"P": P, "U": U, "phase": phase, "fraction": fraction, "region": region}
    return params, atoms, fields, step, cfg["output_file"]


def _prepare_params(args, cfg, params, fe_ids, atoms, lims):
    """Write {output_file}_params.h5 and the output skeleton (cli.py:105-147); returns "prepared"."""
    full = (params.nx, params.ny, params.nz)
    with open_file(cfg["platt_file"], "r") as fh:
        step = int(fh.attrs["Timestep"])
    print(f"\nTimestep {step}:")
    dc = params.reference_strain(np.diag(atoms.aC), params.temperature(step))
    print("\nApplying domain slice:")
    limits = params.apply_slice(*lims)
    params.setup()
    G, qC = reciprocal_centre(params, atoms)
    phase_of = np.full(len(params.names), -1)
    phase_of[fe_ids] = np.arange(len(fe_ids))
    out = Path(f"{cfg['output_file']}.h5").resolve()
    params_file = Path(f"{cfg['output_file']}_params.h5").resolve()
    write_diffraction_axes(cfg["output_file"], *q_axes(params, qC), step)
    resolve = lambda v: Path(v).resolve() if v is not None else None
    write_mgpu_params(params_file, full, limits, (params.dx, params.dy, params.dz), G, qC, np.linalg.det(atoms.aC),
                      atoms.xAtom @ atoms.aC, atoms.fAtom, np.stack([atoms.su1, atoms.su2, atoms.su3]) / OS0,
                      phase_of, dc, resolve(cfg["platt_file"]), resolve(cfg["displacement_file"]),
                      (Path(args.system).parent / "data" / "materialField_0.h5").resolve(),
                      resolve(cfg["phase_file"]), resolve(cfg["region_file"]), out)
    print("\nRun it on one GPU with:")
    print(f"  python -m qpop_diffraction.params_run {params_file}")
    return "prepared"


def _prepare_on_device(args, cfg, params, fe_ids, atoms, full, dz, sl, lims):
    """prepare_run with the fields built on the GPU from the raw cropped components."""
    from .ingest import ingest, read_components, read_material_columns
    real = np.float32 if args.precision == 32 else np.float64
    print(f"Loading P_Lattice file: {cfg['platt_file']}")
    P_raw = read_components(cfg["platt_file"], ("P_Lattice", "PLatt"), *full, sl)
    with open_file(cfg["platt_file"], "r") as fh:
        step = int(fh.attrs["Timestep"])
    U_raw = None
    if cfg["displacement_file"] is not None:
        print(f"Loading Displacement file: {cfg['displacement_file']}")
        U_raw = read_components(cfg["displacement_file"], ("Displacement",), *full, sl)
    else:
        print("No displacement_file provided — using zero displacement field.")
    mat_file = Path(args.system).parent / "data" / "materialField_0.h5"
    print(f"Loading material field: {mat_file}")
    mat = read_material_columns(mat_file, *full, sl)
    phase_of = np.full(len(params.names), -1, dtype=np.int8)
    phase_of[fe_ids] = np.arange(len(fe_ids), dtype=np.int8)
    fraction = None
    if cfg["phase_file"] is not None:
        print(f"Loading phase fractions: {cfg['phase_file']}")
        with open_file(cfg["phase_file"], "r") as fh:
            fraction = np.stack([read_field(fh, f"phase_{p + 1}", *full, sl, real) for p in range(params.nPhase)])
    region = None
    if cfg["region_file"] is not None:
        print(f"Loading scattering weight: {cfg['region_file']}")
        with open_file(cfg["region_file"], "r") as fh:
            region = read_field(fh, "region", *full, sl, real)
    print(f"\nTimestep {step}:")
    dc = params.reference_strain(np.diag(atoms.aC), params.temperature(step))
    P, U, phase = ingest(P_raw, U_raw, mat, dc, phase_of, sl[2].start, dz, args.precision)
    print("\nApplying domain slice:")
    params.apply_slice(*lims)
    params.setup()
    fields = {"P": P, "U": U, "phase": phase, "fraction": fraction, "region": region}
    return params, atoms, fields, step, cfg["output_file"]


def run_z_fastest(nx, nz):
    """Whether the CLI runs its x-fastest fields z fastest: for a run one cell thick along z
    (a 2-D run: pass 0 then has nothing to transform, so the FP32 rounding of its spectra,
    which only the later transforms spread out, is that of the source alone, over the whole
    x-y plane), and when no compile-time source kernel has the x length and one has the z
    length (the source pass then transforms z rows; the generic kernel on long x rows forms
    each cell's source once per field).  QPD_CLI_LAYOUT=x or z forces the layout."""
    forced = os.environ.get("QPD_CLI_LAYOUT", "")
    if forced in ("x", "z"):
        return forced == "z"
    if nz == 1 and nx > 1:
        return True
    kernels = set(fast_lengths("source")) | set(fast_lengths("source_wave"))
    return nx not in kernels and nz in kernels


def intensity_of_fields(engine, params, fields, precision, log=print):
    """The fftshifted intensity (device float64, x fastest) of a run's fields in file order.

    fields: dict with P, U, phase, fraction, region (host arrays or device tensors, x fastest);
    it is emptied, so that the only references to the x-fastest copies are this function's.
    When ``run_z_fastest`` chooses the z-fastest layout the fields are reversed on the device,
    the x-fastest copies released, and the intensity reversed back (peak HBM: the fields twice,
    briefly, then the z-fastest fields, the plan's workspace and two intensities).
    """
    dims, axis = (params.nx, params.ny, params.nz), (0, 1, 2)
    vol = Volume(dims, axis, fields.pop("P"), fields.pop("U"), fields.pop("phase"), fields.pop("fraction", None),
                 fields.pop("region", None), precision)
    fields.clear()
    if not run_z_fastest(params.nx, params.nz):
        return engine.intensity(vol)
    # the fields reversed to z fastest on the device, and the intensity back to x fastest
    rev = lambda t: None if t is None else reverse_axes(t, dims)
    zvol = Volume(dims[::-1], axis[::-1], rev(vol.P), rev(vol.U), rev(vol.phase), rev(vol.fraction), rev(vol.region),
                  precision)
    del vol
    log(f"  fields run z fastest ({params.nx} x {params.ny} x {params.nz} cells)")
    Iz = engine.intensity(zvol)
    del zvol
    return reverse_axes(Iz, dims[::-1])


def main(argv=None) -> int:
    """Run one diffraction calculation; 0 on success, 1 when parameter.atom.in is unusable."""
    args = _parse_args(argv)
    if args.precision is None:
        args.precision = default_precision()
    prepared = prepare_run(args)
    if prepared is None:
        return 1
    if prepared == "prepared":
        return 0
    params, atoms, f, step, out_base = prepared

    print(f"\nComputing diffraction (MI355X, complex FP{args.precision})")
    t0 = time.perf_counter()
    engine = Engine(params, atoms)
    I = intensity_of_fields(engine, params, f, args.precision)
    print(f"  GPU path {time.perf_counter() - t0:.3f} s for {params.n} cells")
    qC = engine.tables[4]
    # the intensity streams from HBM into the file slab by slab
    dt = write_diffraction_streamed(out_base, I, *q_axes(params, qC), step)
    print(f"  wrote {out_base}.h5 in {dt:.3f} s")
    print("\nDiffraction simulation completed.")
    return 0


if __name__ == "__main__":
    sys.exit(main())
