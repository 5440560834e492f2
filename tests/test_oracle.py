"""The FP64 oracle against the reference's golden vectors and closed forms (CPU).

The oracle (oracle/qpd_oracle.py) is the checker of every GPU test, so it is
pinned here first:
  * golden vectors written by the reference's own NumPy path (tests/golden/),
    through the reference CLI's pre-processing restated in tests/helpers.py;
  * the brute-force DFT against numpy.fft (the transform the oracle calls);
  * the closed-form answers of the reference's analytic tests: the single-domain
    structure factor (reference tests/test_single_domain.py:51-58), the stripe
    satellites (tests/test_planted.py:161-173) and the weight scaling
    (tests/test_planted.py:206-219).
A planted mistake (dropped first-order term, wrong sign of the k-space phase)
must fail the golden check; test_planted_mistakes_fail shows it does.
"""

import math

import numpy as np
import pytest

import golden_cases
from conftest import copy_example, load_golden
from helpers import golden_check, oracle_inputs, oracle_run
from qpd_oracle import dft_direct, diffraction_fp64, q_axes, reciprocal_centre
from qpop_diffraction.h5lite import open_file


@pytest.mark.parametrize("name", sorted(golden_cases.CASES))
def test_oracle_matches_reference_golden(tmp_path, name):
    run = golden_cases.CASES[name](open_file, str(tmp_path))
    g = load_golden(name)
    np.testing.assert_allclose(golden_cases.synth.checksum(run), g["checksum"], rtol=1e-10)
    I, q = oracle_run(tmp_path)
    golden_check(I, g, rtol_top=1e-9, rtol_rel=1e-6)
    for a, b in zip(q, (g["q_x"], g["q_y"], g["q_z"])):
        np.testing.assert_allclose(a, b, rtol=1e-14)


def test_oracle_matches_reference_example(tmp_path):
    path = copy_example("2_two_variant_twin", tmp_path / "ex")
    I, q = oracle_run(path)
    golden_check(I, load_golden("example_2_two_variant_twin"), rtol_top=1e-9, rtol_rel=1e-6)


def test_planted_mistakes_fail(tmp_path, monkeypatch):
    """A dropped first-order term or a flipped phase sign in the oracle fails the golden check."""
    import qpd_oracle
    golden_cases.case_bulk(open_file, str(tmp_path))
    g = load_golden("planted_bulk")
    params, atoms, oPhase, oStruc, u, region, _ = oracle_inputs(tmp_path)

    # 1) no first-order term: u enters only through the phase
    real_fftn = np.fft.fftn

    def no_fu(a, *args, **kw):
        out = real_fftn(a, *args, **kw)
        return out * 0 if kw.get("axes") == (1, 2, 3) else out

    monkeypatch.setattr(qpd_oracle.np.fft, "fftn", no_fu)
    I, *_ = diffraction_fp64(params, atoms, oPhase, oStruc, u, region)
    monkeypatch.setattr(qpd_oracle.np.fft, "fftn", real_fftn)
    with pytest.raises(AssertionError):
        golden_check(I.transpose(2, 1, 0), g, rtol_top=1e-9, rtol_rel=1e-6)

    # 2) u with the wrong sign
    I, *_ = diffraction_fp64(params, atoms, oPhase, oStruc, -u, region)
    with pytest.raises(AssertionError):
        golden_check(I.transpose(2, 1, 0), g, rtol_top=1e-9, rtol_rel=1e-6)


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 3, 5), (8, 13, 4), (17, 7, 11), (6, 1, 9)])
def test_direct_dft_matches_numpy(shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    np.testing.assert_allclose(dft_direct(x), np.fft.fftn(x), rtol=1e-12, atol=1e-11)


def test_reciprocal_centre_known_answers():
    # cubic 0.4 nm cell around q00 near 002; strainAvg e shifts qC to G (1 - e) in-plane components
    aC = np.diag([0.4, 0.4, 0.4])
    G, qC = reciprocal_centre(aC, (0.1, -0.05, 31.0), np.zeros(6))
    np.testing.assert_allclose(G, [0.0, 0.0, 2 * math.pi * 2 / 0.4], atol=1e-12)
    np.testing.assert_allclose(qC, G)
    G, qC = reciprocal_centre(aC, (15.0, 0.0, 0.0), [0.01, 0.0, 0.0, 0.0, 0.0, 0.0])
    np.testing.assert_allclose(qC, [G[0] * 0.99, 0.0, 0.0], rtol=1e-14)


def test_q_axes_are_shifted_fft_frequencies():
    qC = np.array([0.1, 0.2, 31.0])
    qx, qy, qz = q_axes((4, 5, 1), (0.5, 0.5, 0.5), qC)
    np.testing.assert_allclose(qx, qC[0] + 2 * np.pi * np.array([-2, -1, 0, 1]) / (4 * 0.5))
    np.testing.assert_allclose(qy, qC[1] + 2 * np.pi * np.array([-2, -1, 0, 1, 2]) / (5 * 0.5))
    np.testing.assert_allclose(qz, [31.0])


def _single_domain(tmp_path, P, hkl):
    nx, ny, nz = 12, 10, 8
    fields = np.zeros((3, nz, ny, nx))
    for d in range(3):
        fields[d] = P[d]
    run = golden_cases.plant_run(fields, np.zeros_like(fields), atom=golden_cases.plant_atom(hkl))
    golden_cases.synth.write_run(str(tmp_path), run, open_file, layers=([1], [nz]))
    return oracle_run(tmp_path)


@pytest.mark.parametrize("P,hkl", [((0, 0, 0.75), (0, 0, 2)), ((0.75, 0, 0), (2, 0, 0)), ((0, -0.75, 0), (1, 1, 1))])
def test_oracle_structure_factor(tmp_path, P, hkl):
    """I(G) = |sum_a f_a exp(-i G.(R_a + d_a))|^2 / V^2 (reference tests/test_single_domain.py:51-58)."""
    I, q = _single_domain(tmp_path, P, hkl)
    A = golden_cases.PLANT_A
    G = 2 * np.pi * np.array(hkl, dtype=float) / A
    r = np.array(golden_cases.PLANT_POS) * A + golden_cases.PLANT_SHIFT * np.asarray(P) / 1e-9
    F = np.sum(np.array(golden_cases.PLANT_FORM) * np.exp(-1j * (r @ G)))
    # the peak voxel is q = G (k = 0): index n//2 of every shifted axis
    peak = I[I.shape[0] // 2, I.shape[1] // 2, I.shape[2] // 2]
    np.testing.assert_allclose(peak, abs(F) ** 2 / A ** 6, rtol=1e-9)
    assert peak == I.max()


def test_oracle_weight_scaling(tmp_path):
    """Weights w = phase fraction x region scale the amplitude linearly (reference tests/test_planted.py:206-219)."""
    nx, ny, nz = golden_cases.PLANT_N
    P = np.zeros((3, nz, ny, nx))
    P[2] = 0.5
    base = golden_cases.plant_run(P, np.zeros_like(P))
    (tmp_path / "plain").mkdir()
    golden_cases.synth.write_run(str(tmp_path / "plain"), base, open_file, layers=([1], [nz]))
    (tmp_path / "w").mkdir()
    golden_cases.synth.write_run(str(tmp_path / "w"), base, open_file, layers=([1], [nz]),
                                 phase=np.full((1, nz, ny, nx), 0.6), region=np.full((nz, ny, nx), 0.5))
    I0, _ = oracle_run(tmp_path / "plain")
    I1, _ = oracle_run(tmp_path / "w")
    mask = I0 > 1e-20 * I0.max()  # voxels above the FP64 round-off floor
    np.testing.assert_allclose(I1[mask], 0.09 * I0[mask], rtol=1e-6)


def test_rsm_map_closed_forms():
    """The oracle's reciprocal-space map on a rectilinear q grid: a default-bin `sum` map is the
    intensity summed over the third axis, `mean` that sum over the axis length, and any binning
    keeps the total intensity."""
    from qpd_oracle import rsm_map
    rng = np.random.default_rng(4)
    I = rng.random((6, 5, 4)) ** 3
    q = [np.linspace(-1, 1, 6) + 3.0, np.linspace(-2, 2, 5), np.linspace(30.0, 34.0, 4)]
    for a, b in ((0, 1), (0, 2), (1, 2)):
        c = 3 - a - b
        ca, cb, m = rsm_map(I, q, a, b)
        np.testing.assert_allclose(ca, q[a])
        np.testing.assert_allclose(cb, q[b])
        np.testing.assert_allclose(m, I.sum(axis=c), rtol=1e-13)
        _, _, mean = rsm_map(I, q, a, b, aggregation="mean")
        np.testing.assert_allclose(mean, I.sum(axis=c) / I.shape[c], rtol=1e-13)
        _, _, coarse = rsm_map(I, q, a, b, n_bins=3)
        assert coarse.shape == (3, 3)
        np.testing.assert_allclose(coarse.sum(), I.sum(), rtol=1e-13)
