"""a1/a2 stripe domains in a film: closed-form thickness fringes and domain satellites.

The case of the reference's tests/test_a1a2_domains.py:50-80, restated from the kinematic
theory.  In-plane P along x (a1 cells, (i + j) mod p < p/2) or along y (a2) fills a film
t cells thick on a non-scattering substrate, with non-scattering cells above; the
reflection is 2 0 0.  With F1, F2 the single-domain structure factors of the variants,
F(q) = sum_a f_a exp(-i q.(R_a + d_a)), d_a the shift of atom a under the variant's P:
  * along q_z through G the film is a uniform slab whose in-plane mean is (F1 + F2) / 2:
      I(G + k_z) / I(G) = sinc^2(k_z t / 2 pi) |F(G + k_z)|^2 / |F(G)|^2   (k_z in 1/cell);
    the shifts are in-plane, so k.d = 0 on the rod and the first-order expansion is exact;
  * the first satellites of the square wave sit at k1 = (2 pi / p)(1, 1, 0):
      I(G + k1) / I(G) = S(k1)^2 |c1|^2 |F1(G + k1) - F2(G + k1)|^2 / |F(G)|^2,
    S the voxel shape factor sinc(k_x / 2 pi) sinc(k_y / 2 pi), c1 the square wave's first
    Fourier coefficient; the first-order expansion in k.d costs about 1e-2 here.  Nothing
    appears along (1, -1, 0), and the even orders of an equal-width wave vanish.
The oracle is held to these on the CPU; the HIP path through the CLI on the GPU, FP64 to
the same bounds and FP32 to the bulk bound (1e-5 of the peak).
"""

import numpy as np
import pytest

import golden_cases
from helpers import oracle_run, run_cli
from qpop_diffraction.h5lite import open_file

synth = golden_cases.synth
A = golden_cases.PLANT_A  # nm
NX, NY, NZ = 32, 48, 32
SUBSTRATE = 8
HKL = (2, 0, 0)
G = 2 * np.pi * np.array(HKL, dtype=float) / A
PS = 0.6  # C/m^2


def structure_factor(q, P):
    """sum_a f_a exp(-i q.(R_a + d_a)) of a uniform domain (positions in nm)."""
    r = np.array(golden_cases.PLANT_POS) * A + golden_cases.PLANT_SHIFT * np.asarray(P) / 1e-9
    return np.sum(np.array(golden_cases.PLANT_FORM) * np.exp(-1j * (r @ q)))


F1 = lambda q: structure_factor(q, (PS, 0.0, 0.0))
F2 = lambda q: structure_factor(q, (0.0, PS, 0.0))


def stripes_run(path, t, p):
    i = np.arange(NX)[None, None, :]
    j = np.arange(NY)[None, :, None]
    a1 = np.broadcast_to(((i + j) % p) < p // 2, (NZ, NY, NX))
    P = np.zeros((3, NZ, NY, NX))
    P[0] = PS * a1
    P[1] = PS * ~a1
    types, sizes = [0, 1, 0], [SUBSTRATE, t, NZ - SUBSTRATE - t]
    mat = np.repeat(np.repeat(types, sizes), NY * NX).reshape(NZ, NY, NX).astype(np.int32)
    mats = [synth.material("Substrate", False, A * 1e-9), synth.material("Film", True, A * 1e-9)]
    run = golden_cases.plant_run(P, np.zeros_like(P), mat=mat, materials=mats, atom=golden_cases.plant_atom(HKL))
    synth.write_run(str(path), run, open_file, geometry="Thin-film", layers=(types, sizes))


def at(I, q, qv):
    """Voxel of I (nz, ny, nx) at the grid point q = qv."""
    idx = [int(np.argmin(np.abs(np.asarray(q[d]) - qv[d]))) for d in range(3)]
    for d in range(3):
        assert abs(q[d][idx[d]] - qv[d]) < 1e-9, (d, q[d][idx[d]], qv[d])
    return I[idx[2], idx[1], idx[0]]


def check(I, q, t, p, floor, rtol_fringe, rtol_sat, atol_rel=0.0):
    I0 = at(I, q, G)
    F = lambda qq: (F1(qq) + F2(qq)) / 2
    # thickness fringes along the rod through G
    for kz in 2 * np.pi * np.fft.fftfreq(NZ, 1.0):
        qz = G + np.array([0.0, 0.0, kz])
        expected = np.sinc(kz * t / (2 * np.pi)) ** 2 * abs(F(qz)) ** 2 / abs(F(G)) ** 2
        got = at(I, q, qz) / I0
        if expected > 100 * floor:
            assert abs(got - expected) <= rtol_fringe * expected + atol_rel, (kz, got, expected)
        elif expected < 1e-20:
            assert got < floor, (kz, got)
    # first-order satellites, nothing along (1, -1, 0), no even orders
    kappa = 2 * np.pi / p
    k1 = kappa * np.array([1.0, 1.0, 0.0])
    c1 = np.mean(np.where(np.arange(p) < p // 2, 1.0, 0.0) * np.exp(-2j * np.pi * np.arange(p) / p))
    S = np.sinc(kappa / (2 * np.pi)) ** 2
    for sgn in (1, -1):
        kk = sgn * k1
        expected = S ** 2 * abs(c1) ** 2 * abs(F1(G + kk) - F2(G + kk)) ** 2 / abs(F(G)) ** 2
        got = at(I, q, G + kk) / I0
        assert abs(got - expected) <= rtol_sat * expected, (sgn, got, expected)
    assert at(I, q, G + kappa * np.array([1.0, -1.0, 0.0])) / I0 < floor
    assert at(I, q, G + 2 * k1) / I0 < floor
    assert I0 == I.max()


@pytest.mark.parametrize("t,p", [(8, 8), (16, 16)])
def test_oracle_a1a2_domains(tmp_path, t, p):
    stripes_run(tmp_path, t, p)
    I, q = oracle_run(tmp_path)
    check(I, q, t, p, floor=1e-20, rtol_fringe=1e-6, rtol_sat=1e-2)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("t,p", [(8, 8), (8, 16), (16, 8), (16, 16)])
def test_gpu_a1a2_domains(tmp_path, t, p, precision):
    print(f"a1/a2 stripes t={t} p={p} precision {precision}", flush=True)
    stripes_run(tmp_path, t, p)
    I, q = run_cli(tmp_path, precision=precision)
    if precision == 64:
        check(I, q, t, p, floor=1e-20, rtol_fringe=1e-6, rtol_sat=1e-2)
    else:
        check(I, q, t, p, floor=1e-5, rtol_fringe=1e-4, rtol_sat=1e-2, atol_rel=1e-5)
