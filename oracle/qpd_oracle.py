"""FP64 CPU oracle of the kinematic-diffraction path.

TEST INFRASTRUCTURE ONLY.  This module is the checker of the HIP path: only the
tests (tests/), the smoke check of __graft_entry__ and the ``cpu_baseline`` leg of
bench.py import it, and only to compute expected values or to time the CPU
reference.  The product path (qpop_diffraction) never imports or calls it.

It restates, in NumPy complex128, the reference computation
src/python/diffraction.py:36-113 (reciprocal_centre, wavevectors, q_axes,
diffraction), line for line in its arithmetic: per atom and phase the displaced
atom positions u_a = U + su . P (diffraction.py:101), the weighted source
fE, fU (:102-104), four numpy.fft.fftn transforms divided by n (:105-106), the
k-space phase and first-order term (:107-109), and I = fftshift(max(|S amp|^2,
1e-50)) (:111).  ``dft_direct`` is a brute-force O(N^2)-per-axis DFT used to pin
numpy.fft on tiny sizes.

How each function is pinned (tests/test_oracle.py, tests/test_planted_forms.py,
tests/test_a1a2_domains.py):
  * reciprocal_centre, q_axes: closed forms and the reference's known answers
    (tests/test_planted.py:141-194 of the reference);
  * diffraction_fp64: the analytic answers of the reference's tests/test_single_domain.py,
    tests/test_planted.py and tests/test_a1a2_domains.py restated from the kinematic theory
    (single-domain structure factors for six variants at four reflections, cubic variant
    symmetry, peak shifts under a u_z ramp, a spontaneous strain and a film lattice,
    stripe satellites on one axis, thickness fringes and domain satellites of a1/a2 films,
    weight scaling), and golden vectors written by the reference's own NumPy path
    (tests/golden/make_golden.py);
  * dft_direct: the DFT definition, checked against numpy.fft;
  * rsm_map: the reciprocal-space map of src/python/rsm.py:66-113 (bin edges from the
    unique q values or a linspace, numpy.histogram2d of the voxels), pinned by its
    closed form on the rectilinear q grid (a `sum` map with default bins is the
    intensity summed over the third axis, a `mean` map that sum over its length) and
    by conservation of the total intensity for any bins.
"""

from __future__ import annotations

import math

import numpy as np


def reciprocal_centre(aC, q00, strainAvg):
    """G (nearest reciprocal lattice point to q00) and qC = G - eps.G (diffraction.py:36-48)."""
    aC = np.asarray(aC, dtype=np.float64)
    G = np.round(aC @ np.asarray(q00, dtype=np.float64) / (2 * math.pi)) @ (2 * math.pi * np.linalg.inv(aC).T)
    s = np.asarray(strainAvg, dtype=np.float64)
    eps = np.array([[s[0], s[5], s[4]], [s[5], s[1], s[3]], [s[4], s[3], s[2]]])
    return G, G - eps @ G


def wavevectors(shape, spacing):
    """2 pi fftfreq per axis (diffraction.py:51-54)."""
    return [2 * math.pi * np.fft.fftfreq(m, d) for m, d in zip(shape, spacing)]


def q_axes(shape, spacing, qC):
    """fftshift(k) + qC per axis (diffraction.py:57-59)."""
    return tuple(np.fft.fftshift(kd) + qC[d] for d, kd in enumerate(wavevectors(shape, spacing)))


def diffraction_fp64(params, atoms, oPhase, oStruc, u, region):
    """The reference's diffraction() on NumPy (diffraction.py:62-113); returns (I, qx, qy, qz)."""
    shape = (params.nx, params.ny, params.nz)
    spacing = (params.dx, params.dy, params.dz)
    n = params.n
    aC = atoms.aC
    V = np.linalg.det(aC)
    G, qC = reciprocal_centre(aC, atoms.q00, params.strainAvg)

    k = wavevectors(shape, spacing)
    kg = [kd.reshape([-1 if d == i else 1 for i in range(3)]) for d, kd in enumerate(k)]
    S = kg[0] * 0 + 1.0
    for d, h in enumerate(spacing):
        S = S * np.sinc(kg[d] * h / (2 * math.pi))

    P = np.asarray(oStruc, dtype=np.float64).reshape(params.nStruc, n)
    U = np.asarray(u, dtype=np.float64).reshape(3, n)
    w = np.asarray(oPhase, dtype=np.float64).reshape(params.nPhase, n) * np.asarray(region, dtype=np.float64)
    su = np.stack([atoms.su1, atoms.su2, atoms.su3])
    R = atoms.xAtom @ aC

    amp = np.zeros(shape, dtype=complex)
    for a in range(atoms.nAtom):
        fE = np.zeros(n, dtype=complex)
        fU = np.zeros((3, n), dtype=complex)
        for p in range(params.nPhase):
            ua = U + su[:, :, p, a] @ P
            e = atoms.fAtom[p, a] * w[p] * np.exp(-1j * (qC @ ua))
            fE += e
            fU += e * ua
        FE = np.fft.fftn(fE.reshape(shape)) / n
        FU = np.fft.fftn(fU.reshape(3, *shape), axes=(1, 2, 3)) / n
        phase = (np.exp(-1j * (kg[0] + G[0]) * R[a, 0]) * np.exp(-1j * (kg[1] + G[1]) * R[a, 1])
                 * np.exp(-1j * (kg[2] + G[2]) * R[a, 2]) / V)
        amp += phase * (FE - 1j * (kg[0] * FU[0] + kg[1] * FU[1] + kg[2] * FU[2]))

    I = np.fft.fftshift(np.maximum(np.abs(S * amp) ** 2, 1e-50))
    return (I,) + q_axes(shape, spacing, qC)


def dft_direct(x, axes=None):
    """Forward DFT sum_j x_j exp(-2 pi i j k / N) along each axis, by matrix products (tiny sizes only)."""
    y = np.asarray(x, dtype=np.complex128)
    axes = range(y.ndim) if axes is None else axes
    for ax in axes:
        N = y.shape[ax]
        j = np.arange(N)
        M = np.exp(-2j * np.pi * np.outer(j, j) / N)
        y = np.moveaxis(np.tensordot(M, np.moveaxis(y, ax, 0), axes=(1, 0)), 0, ax)
    return y


def relative_to_peak(I, ref):
    """max |I - ref| / max(ref): the parity measure of intensities."""
    return float(np.max(np.abs(np.asarray(I) - np.asarray(ref))) / np.max(ref))


def max_dlog10(I, ref, floor_rel=1e-10):
    """max |log10 I - log10 ref| over voxels with ref >= floor_rel * max(ref)."""
    I = np.asarray(I, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    mask = ref >= floor_rel * ref.max()
    return float(np.max(np.abs(np.log10(np.maximum(I[mask], 1e-300)) - np.log10(ref[mask]))))


def rsm_edges(values, n_bins=None):
    """Bin edges and centres of one q component (rsm.py:66-85)."""
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


def rsm_map(I, q, axis_a, axis_b, n_bins=None, aggregation="sum"):
    """Map of the (nx, ny, nz) intensity on the q_a - q_b plane (rsm.py:88-113), every voxel histogrammed."""
    grids = np.meshgrid(*q, indexing="ij")
    qa, qb = grids[axis_a].ravel(), grids[axis_b].ravel()
    ea, ca = rsm_edges(qa, n_bins)
    eb, cb = rsm_edges(qb, n_bins)
    total, _, _ = np.histogram2d(qa, qb, bins=(ea, eb), weights=np.asarray(I).ravel())
    if aggregation == "sum":
        return ca, cb, total
    counts, _, _ = np.histogram2d(qa, qb, bins=(ea, eb))
    return ca, cb, np.divide(total, counts, out=np.zeros_like(total), where=counts > 0)
