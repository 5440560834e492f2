"""Grids of any axis length against the FP64 oracle.

numpy.fft (the reference's xp.fft.fftn, diffraction.py:105-106) and HeFFTe transform every
length; the library does too: prime factors above 17 run as direct stages (fft_stage_direct),
rows of more than 4096 points at sixteen points a thread.  The CLI runs a 38 x 23 x 19 grid
(2 x 19, 23, 19); the drop-in volume path runs prime and long axes along each memory dim.
FP64 plans are held to the reference's parity bound (rtol 1e-9, 1e-12 of the peak), FP32
plans to the FP32 bounds (1e-5 of the peak, |dlog10 I| <= 1e-3, same peak voxel).
"""

import numpy as np
import pytest

from helpers import check_fp32, oracle_run, run_cli
from qpop_diffraction import synth
from qpop_diffraction.h5lite import open_file

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [64, 32])
def test_cli_prime_lengths(tmp_path, precision):
    print(f"cli 38x23x19 precision {precision}", flush=True)
    run = synth.two_phase_run(38, 23, 19, n_grains=5, seed=11)
    synth.write_run(str(tmp_path), run, open_file)
    I, q = run_cli(tmp_path, precision=precision)
    ref, qref = oracle_run(tmp_path)
    assert I.shape == (19, 23, 38)
    for a, b in zip(q, qref):
        np.testing.assert_allclose(a, b, rtol=1e-14)
    if precision == 64:
        np.testing.assert_allclose(I, ref, rtol=1e-9, atol=1e-12 * ref.max())
    else:
        check_fp32(I, ref, "cli 38x23x19")


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("dims", [(97, 5, 4), (4, 97, 5), (5, 4, 97), (4001, 2, 3), (3, 4001, 2), (2, 3, 4001),
                                  (5003, 2, 2), (2, 2, 8000)])
def test_volume_prime_and_long_axes(dims, precision):
    """The drop-in volume path (memory dims x fastest) on prime and long axes."""
    import torch
    from qpd_oracle import diffraction_fp64
    from qpop_diffraction import Engine, Volume, workloads
    nx, ny, nz = dims
    print(f"volume {dims} precision {precision}", flush=True)
    w = workloads.two_phase(nx, ny, nz, n_grains=3, seed=6, precision=precision)
    eng = Engine(w.params, w.atoms)
    vol = Volume(w.dims, w.axis, w.P, w.U, w.phase, precision=precision)
    I = eng.intensity(vol).cpu().numpy().reshape(nz, ny, nx)
    to_xyz = lambda t, k: t.double().reshape(k, nz, ny, nx).permute(0, 3, 2, 1).contiguous().cpu().numpy()
    ph = to_xyz(w.phase, 1)[0]
    oPhase = np.stack([(ph == p).astype(np.float64) for p in range(2)]).ravel()
    reg = (ph >= 0).astype(np.float64).ravel()
    ref = diffraction_fp64(w.params, w.atoms, oPhase, to_xyz(w.P, 3).ravel(), to_xyz(w.U, 3).ravel(),
                           reg)[0].transpose(2, 1, 0)
    torch.cuda.synchronize()
    if precision == 64:
        np.testing.assert_allclose(I, ref, rtol=1e-9, atol=1e-12 * ref.max())
    else:
        check_fp32(I, ref, f"volume {dims}")


@pytest.mark.parametrize("precision", [64, 32])
def test_generic_groups_pass_every_schedule(precision):
    """Pass 1 over source groups on a 97-point axis (no compile-time kernel): every generic
    schedule against the oracle."""
    import torch
    from qpd_oracle import diffraction_fp64
    from qpop_diffraction import Engine, Volume, workloads
    nx, ny, nz = 6, 97, 5
    w = workloads.two_phase(nx, ny, nz, n_grains=3, seed=8, precision=precision)
    eng = Engine(w.params, w.atoms, order="memory")
    vol = Volume(w.dims, w.axis, w.P, w.U, w.phase, precision=precision)
    plan = eng.plan(vol.dims, vol.axis, precision)
    names = plan.schedules(1)
    groups = [i for i, nm in enumerate(names) if " groups" in nm]
    assert len(groups) >= 2, names
    to_xyz = lambda t, k: t.double().reshape(k, nz, ny, nx).permute(0, 3, 2, 1).contiguous().cpu().numpy()
    ph = to_xyz(w.phase, 1)[0]
    oPhase = np.stack([(ph == p).astype(np.float64) for p in range(2)]).ravel()
    ref = diffraction_fp64(w.params, w.atoms, oPhase, to_xyz(w.P, 3).ravel(), to_xyz(w.U, 3).ravel(),
                           (ph >= 0).astype(np.float64).ravel())[0].transpose(2, 1, 0)
    for i in groups:
        print(f"precision {precision} pass 1 schedule {names[i]}", flush=True)
        plan.set_schedule(1, i)
        I = eng.intensity(vol).cpu().numpy().reshape(nz, ny, nx)
        torch.cuda.synchronize()
        if precision == 64:
            np.testing.assert_allclose(I, ref, rtol=1e-9, atol=1e-12 * ref.max())
        else:
            check_fp32(I, ref, names[i])
