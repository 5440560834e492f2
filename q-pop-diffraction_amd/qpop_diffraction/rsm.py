"""Reciprocal-space maps: the intensity projected onto the q_x-q_y, q_x-q_z and q_y-q_z planes
(reference: src/python/rsm.py:21-175, console script ``qpop-diffraction-rsm``).

The reference histograms every voxel of the 3-D intensity with numpy.histogram2d.  On the
rectilinear q grid every voxel of a column along the third axis falls in the same bin, so
the map is the histogram of the two-axis *sum* of the intensity, a small 2-D array.  Those
sums run on the GPU (``qpd_project``: one pass per axis over the intensity in HBM, fixed
summation order), so a map of a volume that is still on the device costs a few
milliseconds and no device-to-host copy of the volume; only the 2-D histogram runs on the
host, with the reference's bin edges.  Maps are written as ``{pair}.npz`` with the
reference's keys, and as ``{pair}.png`` when matplotlib is available.
"""

from __future__ import annotations

import argparse
from pathlib import Path

import numpy as np

from ._native import project

PAIR_DEFINITIONS = ((0, 1, "qx_qy"), (0, 2, "qx_qz"), (1, 2, "qy_qz"))
AXIS_LABELS = ("Qx (nm$^{-1}$)", "Qy (nm$^{-1}$)", "Qz (nm$^{-1}$)")


def build_edges(values, n_bins=None):
    """Bin edges and centres of one q component, as rsm.py:66-85 builds them."""
    values = np.asarray(values, dtype=np.float64)
    if n_bins is not None:
        edges = np.linspace(values.min(), values.max(), n_bins + 1)
        return edges, 0.5 * (edges[:-1] + edges[1:])
    unique = np.unique(np.round(values, decimals=10))
    if unique.size == 1:
        return np.array([unique[0] - 0.5, unique[0] + 0.5]), unique.copy()
    mid = 0.5 * (unique[:-1] + unique[1:])
    first = unique[0] - 0.5 * (unique[1] - unique[0])
    last = unique[-1] + 0.5 * (unique[-1] - unique[-2])
    return np.concatenate(([first], mid, [last])), unique


def axis_sums(I, dims, axis):
    """Device sums of an intensity over each physical axis.

    I: device float64 tensor in a grid layout (dims[0] fastest, axis[d] the physical axis of
    memory dim d).  Returns {c: host array (n_a, n_b)}: the sum over physical axis c, indexed
    by the two other physical axes in increasing order.
    """
    sums = project(I.reshape(-1), dims)
    out = {}
    for d, S in sums.items():
        others = [e for e in range(3) if e != d]           # memory dims left, slower one first in S
        phys = [axis[others[1]], axis[others[0]]]          # physical axes of S's rows, columns
        arr = S.cpu().numpy()
        out[axis[d]] = arr if phys[0] < phys[1] else arr.T
    return out


def _bins(values, edges):
    """Bin of each value as numpy.histogramdd assigns it (right-open bins, the last one closed;
    0 and len(edges) mark values below and above the edges)."""
    idx = np.searchsorted(edges, values, side="right")
    idx[values == edges[-1]] -= 1
    return idx


def _histogram2d_grid(qa_1d, qb_1d, ea, eb, weights):
    """numpy.histogram2d of the grid points (qa_1d[i], qb_1d[j]) in (i, j) C order with weights
    (n_a, n_b): the same bins and the same order of additions (one bincount over the points),
    with the binning done on the 1-D axes."""
    ia, ib = _bins(qa_1d, ea), _bins(qb_1d, eb)
    na, nb = len(ea) + 1, len(eb) + 1
    flat = (ia[:, None] * nb + ib[None, :]).ravel()
    hist = np.bincount(flat, weights=None if weights is None else np.asarray(weights, np.float64).ravel(),
                       minlength=na * nb).reshape(na, nb)
    return hist[1:-1, 1:-1].astype(np.float64)


def maps_from_sums(sums, q, n_bins=None, aggregation="sum"):
    """The three maps from the axis sums and the 1-D q axes (rsm.py:88-113).

    The reference histograms the meshgrid of two q axes; its edges come from the unique values
    of that meshgrid, which are the unique values of the 1-D axis, and its histogram adds the
    points in meshgrid order, which _histogram2d_grid reproduces from the 1-D axes.
    """
    n = [len(v) for v in q]
    maps = {}
    for a, b, stem in PAIR_DEFINITIONS:
        c = 3 - a - b
        qa_1d, qb_1d = np.asarray(q[a], np.float64), np.asarray(q[b], np.float64)
        ea, ca = build_edges(qa_1d, n_bins)
        eb, cb = build_edges(qb_1d, n_bins)
        total = _histogram2d_grid(qa_1d, qb_1d, ea, eb, sums[c])
        if aggregation == "sum":
            data = total
        else:
            counts = _histogram2d_grid(qa_1d, qb_1d, ea, eb, None)
            counts *= n[c]  # every (a, b) point stands for a whole column along c
            data = np.divide(total, counts, out=np.zeros_like(total), where=counts > 0)
        maps[stem] = (ca, cb, data)
    return maps


def reciprocal_space_maps(I, dims, axis, q, n_bins=None, aggregation="sum"):
    """Maps of a device intensity in any grid layout; q = (q_x, q_y, q_z) 1-D axes."""
    return maps_from_sums(axis_sums(I, dims, axis), q, n_bins, aggregation)


def save_map_data(path, q_a, q_b, data):
    """rsm.py:116-126: the map and its log10 with the q axes, as .npz."""
    np.savez(path, q_a=q_a, q_b=q_b, q_units="nm^-1", intensity=data,
             log10_intensity=np.log10(np.maximum(data, np.finfo(np.float64).tiny)))


def save_map_figure(path, q_a, q_b, data, label_a, label_b, aggregation):
    """rsm.py:129-150 when matplotlib is importable; returns whether a figure was written."""
    try:
        import matplotlib
        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
    except ImportError:
        return False
    fig, ax = plt.subplots(figsize=(8, 6))
    image = ax.pcolormesh(q_a, q_b, np.log10(np.maximum(data, np.finfo(np.float64).tiny)).T, shading="auto",
                          cmap="turbo")
    ax.set_xlabel(label_a)
    ax.set_ylabel(label_b)
    ax.set_title(f"Reciprocal-space map: {label_a} vs {label_b} ({aggregation})")
    fig.colorbar(image, ax=ax).set_label("log10 intensity")
    fig.tight_layout()
    fig.savefig(path, dpi=200)
    plt.close(fig)
    return True


def parse_args(argv=None):
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument("--diffraction", "-d", type=Path, required=True, metavar="HDF5",
                        help="Diffraction HDF5 written by qpop-diffraction (/intensity, /q_x, /q_y, /q_z).")
    parser.add_argument("--output-dir", "-o", type=Path, default=Path("reciprocal_space_maps"),
                        help="Directory for generated maps and figures.")
    parser.add_argument("--bins", type=int, default=None,
                        help="Override the number of bins per axis (default: one per q sample).")
    parser.add_argument("--aggregation", choices=("sum", "mean"), default="sum",
                        help="How to accumulate intensity when several voxels fall into one bin.")
    return parser.parse_args(argv)


def main(argv=None) -> int:
    """qpop-diffraction-rsm (rsm.py:153-175): maps of a diffraction.h5, projected on the GPU."""
    import torch
    from .h5lite import open_file, read_dataset
    args = parse_args(argv)
    with open_file(str(args.diffraction), "r") as fh:
        I = read_dataset(fh, "intensity", dtype=np.float64)  # (nz, ny, nx)
        q = [read_dataset(fh, f"q_{c}", dtype=np.float64) for c in "xyz"]
    print(f"Loaded diffraction HDF5: {args.diffraction}")
    args.output_dir.mkdir(parents=True, exist_ok=True)
    nz, ny, nx = I.shape
    dev = torch.from_numpy(np.ascontiguousarray(I)).to("cuda")
    maps = reciprocal_space_maps(dev, (nx, ny, nz), (0, 1, 2), q, args.bins, args.aggregation)
    for a, b, stem in PAIR_DEFINITIONS:
        ca, cb, data = maps[stem]
        save_map_data(args.output_dir / f"{stem}.npz", ca, cb, data)
        fig = save_map_figure(args.output_dir / f"{stem}.png", ca, cb, data, AXIS_LABELS[a], AXIS_LABELS[b],
                              args.aggregation)
        print(f"Wrote {stem}.npz{' and ' + stem + '.png' if fig else ''} with shape {data.shape} to {args.output_dir}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
