"""Planted single-domain, stripe, strain and film runs against their closed-form answers.

The cases of the reference's tests/test_single_domain.py:60-96 and tests/test_planted.py:141-230,
restated from the kinematic theory and the FerroDyn setting rules.  The PbTiO3 cell of the
planted golden cases (complex form factors, anisotropic P -> shift coefficients) on a
40 x 30 x 20 grid of 1 nm cells: with |G_200| / 2 pi = 5 / nm and e = 0.02, e L |G| / 2 pi is
an integer on every axis, so a uniform strain u = e r is periodic in exp(-i G.u).
  * a uniform P along each of the six variants puts the peak at G with I(G) = |F(G)|^2 / V^2,
    F(G) = sum_a f_a exp(-i G.(R_a + d_a)) computed here; u_z = e z moves the 002 peak to G (1 - e) (u in metres:
    a missing m -> nm conversion would leave it at G); a spontaneous strain e along an axis
    moves the peak to G / (1 + e) along it (first order in u: G (1 - e));
  * 180-degree P_z stripes of period L along x give satellites at +-2 pi / L on q_x only;
  * the film's own lattice, not the atom file's aC, sets the 002 peak: 4 pi / c_film;
  * the cell is cubic: P along d seen at the 2 0 0 reflection along d is the same for every d,
    a shift perpendicular to G is invisible, and the imaginary form factors break Friedel's law
    (P and -P differ along G);
  * a region zero on half the x range halves the amplitude at G;
  * the host rules (CPU): substrate misfit and reference strain, the temperature ramp, and
    rotated runs rejected.
The oracle is checked on the CPU; the HIP path through the CLI on the GPU, FP64 to the same
bounds and FP32 to the bulk bound where a voxel is compared with the peak.
"""

import json

import numpy as np
import pytest

import golden_cases
from helpers import oracle_run, run_cli
from qpop_diffraction.h5lite import open_file
from qpop_diffraction.size_context import SizeContext

synth = golden_cases.synth
A = golden_cases.PLANT_A  # nm
NX, NY, NZ = 40, 30, 20
DX = 1e-9  # m
PS = 0.5
VARIANTS = {"c+": (0, 0, 1), "c-": (0, 0, -1), "a1+": (1, 0, 0), "a1-": (-1, 0, 0), "a2+": (0, 1, 0), "a2-": (0, -1, 0)}


def zeros():
    return np.zeros((3, NZ, NY, NX))


def write(path, P, u, hkl=(0, 0, 2), film_a=A * 1e-9, region=None):
    mats = [synth.material("Substrate", False, A * 1e-9), synth.material("Film", True, film_a)]
    run = golden_cases.plant_run(P, u, materials=mats, atom=golden_cases.plant_atom(hkl))
    run["spacing"] = DX
    synth.write_run(str(path), run, open_file, layers=([1], [NZ]), region=region)


def reciprocal(hkl):
    return 2 * np.pi * np.array(hkl, dtype=float) / A


def peak(I, q):
    k = np.unravel_index(np.argmax(I), I.shape)
    return np.array([q[0][k[2]], q[1][k[1]], q[2][k[0]]])


def at(I, q, qv):
    idx = [int(np.argmin(np.abs(np.asarray(q[d]) - qv[d]))) for d in range(3)]
    for d in range(3):
        assert abs(q[d][idx[d]] - qv[d]) < 1e-9, (d, q[d][idx[d]], qv[d])
    return I[idx[2], idx[1], idx[0]]


RUNNERS = [pytest.param("oracle", id="oracle"),
           pytest.param(64, id="gpu64", marks=pytest.mark.gpu),
           pytest.param(32, id="gpu32", marks=pytest.mark.gpu)]


def compute(path, how):
    return oracle_run(path) if how == "oracle" else run_cli(path, precision=how)


def floor(how):
    """Relative level below which a voxel counts as empty: FP64 round-off or the FP32 bulk bound."""
    return 1e-5 if how == 32 else 1e-12


def structure_factor(q, P):
    """sum_a f_a exp(-i q.(R_a + d_a)) of a uniform domain, positions in nm, P in C/m^2."""
    r = np.array(golden_cases.PLANT_POS) * A + golden_cases.PLANT_SHIFT * np.asarray(P) / 1e-9
    return np.sum(np.array(golden_cases.PLANT_FORM) * np.exp(-1j * (r @ q)))


@pytest.mark.parametrize("how", RUNNERS)
@pytest.mark.parametrize("hkl", [(0, 0, 2), (2, 0, 0), (0, 2, 0), (1, 1, 1)], ids=lambda h: "".join(map(str, h)))
@pytest.mark.parametrize("variant", list(VARIANTS))
def test_structure_factor(tmp_path, variant, hkl, how):
    """I(G) = |sum_a f_a exp(-i G.(R_a + d_a))|^2 / V^2 for a uniform domain, peak at G."""
    P = 0.75 * np.array(VARIANTS[variant], dtype=float)
    Pf = zeros()
    for d in range(3):
        Pf[d] = P[d]
    write(tmp_path, Pf, zeros(), hkl=hkl)
    I, q = compute(tmp_path, how)
    G = reciprocal(hkl)
    np.testing.assert_allclose(peak(I, q), G, atol=1e-9)
    np.testing.assert_allclose(at(I, q, G), abs(structure_factor(G, P)) ** 2 / A ** 6, rtol=2e-5 if how == 32 else 1e-9)


@pytest.mark.parametrize("how", RUNNERS)
def test_uniform_p_peaks_at_g(tmp_path, how):
    P = zeros()
    P[2] = PS
    write(tmp_path, P, zeros())
    I, q = compute(tmp_path, how)
    np.testing.assert_allclose(peak(I, q), reciprocal((0, 0, 2)), atol=1e-9)


@pytest.mark.parametrize("how", RUNNERS)
def test_uz_ramp_shifts_peak(tmp_path, how):
    e = 0.02
    u = zeros()
    u[2] = e * DX * np.arange(NZ)[:, None, None]
    write(tmp_path, zeros(), u)
    I, q = compute(tmp_path, how)
    dqz = 2 * np.pi / (NZ * DX * 1e9)
    assert abs(peak(I, q)[2] - reciprocal((0, 0, 2))[2] * (1 - e)) < dqz


@pytest.mark.parametrize("how", RUNNERS)
def test_stripe_satellites_along_x_only(tmp_path, how):
    L = 10
    P = zeros()
    P[2] = np.where((np.arange(NX) // (L // 2)) % 2 == 0, PS, -PS)[None, None, :]
    write(tmp_path, P, zeros())
    I, q = compute(tmp_path, how)
    G = reciprocal((0, 0, 2))
    k = 2 * np.pi / (L * DX * 1e9)
    on_x = at(I, q, G + np.array([k, 0.0, 0.0]))
    on_y = at(I, q, G + np.array([0.0, k, 0.0]))  # the same offset on q_y (on the grid: NY / L = 3)
    assert on_x > 1e-3 * I.max()
    assert on_y < floor(how) * I.max()


@pytest.mark.parametrize("how", RUNNERS)
def test_film_lattice_sets_the_peak(tmp_path, how):
    c = 0.392
    write(tmp_path, zeros(), zeros(), film_a=c * 1e-9)
    I, q = compute(tmp_path, how)
    dqz = 2 * np.pi / (NZ * DX * 1e9)
    assert abs(peak(I, q)[2] - 4 * np.pi / c) < dqz


@pytest.mark.parametrize("how", RUNNERS)
def test_half_region_halves_the_amplitude(tmp_path, how):
    P = zeros()
    P[2] = PS
    (tmp_path / "plain").mkdir()
    write(tmp_path / "plain", P, zeros())
    (tmp_path / "half").mkdir()
    write(tmp_path / "half", P, zeros(), region=(np.arange(NX) < NX // 2)[None, None, :] * np.ones((NZ, NY, NX)))
    I0, q = compute(tmp_path / "plain", how)
    I1, _ = compute(tmp_path / "half", how)
    G = reciprocal((0, 0, 2))
    np.testing.assert_allclose(at(I1, q, G), 0.25 * at(I0, q, G), rtol=1e-6 if how != 32 else 1e-5)


@pytest.mark.parametrize("how", [pytest.param("oracle", id="oracle"), pytest.param(64, id="gpu64", marks=pytest.mark.gpu)])
def test_variant_symmetry(tmp_path, how):
    I = {}
    for name, P in list(VARIANTS.items()) + [("0", (0, 0, 0))]:
        for hkl in [(0, 0, 2), (2, 0, 0), (0, 2, 0)]:
            work = tmp_path / f"{name}_{''.join(map(str, hkl))}"
            work.mkdir()
            Pf = zeros()
            for d in range(3):
                Pf[d] = PS * P[d]
            write(work, Pf, zeros(), hkl=hkl)
            out, q = compute(work, how)
            I[name, hkl] = at(out, q, reciprocal(hkl))
    for sign in "+-":
        parallel = [I["c" + sign, (0, 0, 2)], I["a1" + sign, (2, 0, 0)], I["a2" + sign, (0, 2, 0)]]
        np.testing.assert_allclose(parallel, parallel[0], rtol=1e-12)
        perpendicular = [I["a1" + sign, (0, 0, 2)], I["a2" + sign, (0, 0, 2)], I["c" + sign, (2, 0, 0)]]
        np.testing.assert_allclose(perpendicular, I["0", (0, 0, 2)], rtol=1e-12)
    assert abs(I["c+", (0, 0, 2)] / I["c-", (0, 0, 2)] - 1) > 1e-3


@pytest.mark.parametrize("how", RUNNERS)
@pytest.mark.parametrize("variant", ["c+", "a1+", "a2+"])
def test_spontaneous_strain(tmp_path, variant, how):
    e = 0.02
    axis = int(np.argmax(np.abs(VARIANTS[variant])))
    n = (NX, NY, NZ)
    for hkl in [(0, 0, 2), (2, 0, 0), (0, 2, 0)]:
        work = tmp_path / "".join(map(str, hkl))
        work.mkdir()
        P = zeros()
        for d in range(3):
            P[d] = PS * VARIANTS[variant][d]
        u = zeros()
        r = np.arange(n[axis]) * DX
        u[axis] = e * r.reshape([-1 if i == 2 - axis else 1 for i in range(3)])
        write(work, P, u, hkl=hkl)
        I, q = compute(work, how)
        G = reciprocal(hkl)
        expected = G.copy()
        expected[axis] = G[axis] / (1 + e)
        dq = 2 * np.pi / (np.array(n) * DX * 1e9)
        assert np.all(np.abs(peak(I, q) - expected) < dq / 2), (hkl, peak(I, q), expected)


def test_substrate_misfit(tmp_path):
    """Thin film: the in-plane lattice is the substrate's, referred to aC; the substrate is not a phase."""
    mats = [synth.material("Substrate", False, 3.9e-10), synth.material("Film", True, 4.0e-10)]
    run = golden_cases.plant_run(zeros(), zeros(), materials=mats,
                                 mat=np.repeat(np.repeat([0, 1], [8, 12]), NY * NX).reshape(NZ, NY, NX).astype(np.int32))
    synth.write_run(str(tmp_path), run, open_file, geometry="Thin-film", layers=([0, 1], [8, 12]))
    params = SizeContext()
    assert list(params.read_json(str(tmp_path / "system_setting.json"))) == [1]
    dc = params.reference_strain(np.array([0.41, 0.41, 0.41]), 300.0)
    np.testing.assert_allclose(params.strainAvg, [0.39 / 0.41 - 1, 0.39 / 0.41 - 1, 0, 0, 0, 0])
    np.testing.assert_allclose(dc, [0.39 / 0.41 - 1, 0.4 / 0.41 - 1])


def test_temperature_ramp(tmp_path):
    write(tmp_path, zeros(), zeros())
    params = SizeContext()
    params.read_json(str(tmp_path / "system_setting.json"))
    params.setting["SimulationControls"]["Temperature"] = {
        "Initial": 300.0, "Transient": True, "Final": 700.0, "Interval": 4000, "Step": 25}
    assert [params.temperature(s) for s in (0, 3999, 4000, 7500, 67500, 99999)] == [300, 300, 325, 325, 700, 700]


def test_rotated_run_rejected(tmp_path):
    write(tmp_path, zeros(), zeros())
    setting = json.loads((tmp_path / "system_setting.json").read_text())
    setting["Rotation"]["Enabled"] = True
    (tmp_path / "system_setting.json").write_text(json.dumps(setting))
    with pytest.raises(ValueError):
        SizeContext().read_json(str(tmp_path / "system_setting.json"))
