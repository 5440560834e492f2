"""Run a prepared ``{output}_params.h5`` on one GPU.

The reference prepares large runs with ``qpop-diffraction --mgpu`` (cli.py:105-147): it
writes ``{output}.h5`` with the q axes and ``{output}_params.h5`` (io_utils.py:144-183), and
the MPI program ``qpop_diffraction_mgpu`` reads the latter (params.hpp:13-108), reads its slab
of the fields, and fills in ``/intensity`` (qpop_diffraction_mgpu.cu:60-209, hdf5_io.hpp:64-97).
This module is that program for one MI355X: it reads the whole crop, builds the plan inputs
with the GPU ingest (u_scale and dc, phase_of per material from the file), computes the
volume and streams ``/intensity`` into the prepared output file.

    python -m qpop_diffraction.params_run <output>_params.h5 [--precision 32|64]
"""

from __future__ import annotations

import argparse
import sys
import time
from types import SimpleNamespace

import numpy as np

from .cli import intensity_of_fields
from .diffraction import Engine, default_precision
from .h5lite import open_file, read_dataset
from .ingest import ingest, read_components, read_material_columns
from .io_utils import read_field, read_mgpu_params, write_diffraction_streamed
from .size_context import SizeContext


def _params_of(prm):
    """A SizeContext with the crop's grid and spacing (what the plan and the q axes need)."""
    x0, x1, y0, y1, z0, z1 = (int(v) for v in prm["crop"])
    p = SizeContext()
    p.nx, p.ny, p.nz = x1 - x0 + 1, y1 - y0 + 1, z1 - z0 + 1
    dx, dy, dz = (float(v) for v in prm["spacing"])
    p.lx, p.ly, p.lz = p.nx * dx, p.ny * dy, p.nz * dz
    p.nPhase, p.nStruc = prm["fAtom_re"].shape[0], 3
    p.strainAvg = np.zeros(6)
    p.setup()
    p.dx, p.dy, p.dz = dx, dy, dz  # the file's spacing exactly, not nx * dx / nx
    return p


def run(params_file, precision=None, log=print):
    """Compute the prepared run and write its /intensity; returns (seconds, cells)."""
    prm = read_mgpu_params(params_file)
    precision = default_precision() if precision is None else int(precision)
    t0 = time.perf_counter()
    nx_f, ny_f, nz_f = (int(v) for v in prm["grid"])
    x0, x1, y0, y1, z0, z1 = (int(v) for v in prm["crop"])
    sl = (slice(x0, x1 + 1), slice(y0, y1 + 1), slice(z0, z1 + 1))
    P_raw = read_components(prm["platt_file"], (prm["platt_prefix"],), nx_f, ny_f, nz_f, sl)
    U_raw = None
    if prm["displacement_file"]:
        U_raw = read_components(prm["displacement_file"], ("Displacement",), nx_f, ny_f, nz_f, sl)
    mat = read_material_columns(prm["material_file"], nx_f, ny_f, nz_f, sl)
    params = _params_of(prm)
    real = np.float32 if precision == 32 else np.float64
    fraction = region = None
    if prm["phase_file"]:
        with open_file(prm["phase_file"], "r") as fh:
            fraction = np.stack([read_field(fh, f"phase_{p + 1}", nx_f, ny_f, nz_f, sl, real)
                                 for p in range(params.nPhase)])
    if prm["region_file"]:
        with open_file(prm["region_file"], "r") as fh:
            region = read_field(fh, "region", nx_f, ny_f, nz_f, sl, real)
    dz = float(prm["spacing"][2])
    # su in the file already carries 1/OS0, so P is taken as read (params.hpp:13-108)
    P, U, phase = ingest(P_raw, U_raw, mat, prm["dc"], prm["phase_of"].astype(np.int8), z0, dz, precision,
                         p_div=1.0, u_div=1.0 / float(prm["u_scale"][0]))
    f = prm["fAtom_re"] + 1j * prm["fAtom_im"]
    tables = (prm["R"], f, prm["su"], prm["G"], prm["qC"], float(prm["V"][0]))
    engine = Engine.from_tables(params, tables)
    # the layout rule of the command line, so that both routes compute the same volume
    fields = {"P": P, "U": U, "phase": phase, "fraction": fraction, "region": region}
    I = intensity_of_fields(engine, params, fields, precision, log=log)
    out = prm["output_file"]
    base = out[:-3] if out.endswith(".h5") else out
    with open_file(out, "r") as fh:
        q = [read_dataset(fh, f"q_{c}", dtype=np.float64) for c in "xyz"]
        step = int(fh.attrs["Timestep"])
    write_diffraction_streamed(base, I, *q, step, axes=False)
    dt = time.perf_counter() - t0
    log(f"Computed {params.n} cells on one GPU in {dt:.3f} s; wrote /intensity to {out}")
    return dt, params.n


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("params", help="<output>_params.h5 written by qpop-diffraction --mgpu")
    ap.add_argument("--precision", type=int, choices=(32, 64), default=None)
    args = ap.parse_args(argv)
    run(args.params, args.precision)
    return 0


if __name__ == "__main__":
    sys.exit(main())
