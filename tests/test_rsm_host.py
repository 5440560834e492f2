"""The host step of the reciprocal-space maps: binning on the 1-D q axes reproduces
numpy.histogram2d of the meshgrid (the reference's call, rsm.py:88-113) exactly, for the
default one-bin-per-q edges, for linspace edges (--bins), for axes with values that merge
when rounded to 1e-10, and for the mean aggregation."""

import numpy as np
import pytest

from qpop_diffraction.rsm import build_edges, maps_from_sums, PAIR_DEFINITIONS


def reference_maps(sums, q, n_bins, aggregation):
    out = {}
    n = [len(v) for v in q]
    for a, b, stem in PAIR_DEFINITIONS:
        c = 3 - a - b
        qa, qb = np.meshgrid(q[a], q[b], indexing="ij")
        ea, ca = build_edges(qa.ravel(), n_bins)
        eb, cb = build_edges(qb.ravel(), n_bins)
        total, _, _ = np.histogram2d(qa.ravel(), qb.ravel(), bins=(ea, eb), weights=sums[c].ravel())
        if aggregation == "sum":
            data = total
        else:
            counts, _, _ = np.histogram2d(qa.ravel(), qb.ravel(), bins=(ea, eb))
            counts *= n[c]
            data = np.divide(total, counts, out=np.zeros_like(total), where=counts > 0)
        out[stem] = (ca, cb, data)
    return out


@pytest.mark.parametrize("n_bins", [None, 7, 33])
@pytest.mark.parametrize("aggregation", ["sum", "mean"])
@pytest.mark.parametrize("dup", [False, True])
def test_grid_binning_equals_histogram2d(n_bins, aggregation, dup):
    rng = np.random.default_rng(5)
    n = (24, 17, 10)
    q = [np.fft.fftshift(np.fft.fftfreq(m, 0.4)) * 2 * np.pi + 3.0 * k for k, m in enumerate(n)]
    if dup:
        q[1] = q[1].copy()
        q[1][5] = q[1][4] + 1e-12  # merges with its neighbour after rounding to 1e-10
    sums = {c: rng.random((n[a], n[b])) * 10.0 ** rng.uniform(-8, 3, (n[a], n[b]))
            for a, b, c in ((1, 2, 0), (0, 2, 1), (0, 1, 2))}
    got = maps_from_sums(sums, q, n_bins, aggregation)
    ref = reference_maps(sums, q, n_bins, aggregation)
    for stem in ref:
        for x, y in zip(got[stem], ref[stem]):
            np.testing.assert_array_equal(x, y)
