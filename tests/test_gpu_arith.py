"""Pass order and arithmetic of FP32 plans, against the FP64 oracle.

An FP32 plan transforms the contiguous dim first; its arithmetic is
  * fp32:  complex FP32 transforms and spectra (explicit only: an FP32 transform of a line
           holding a strong peak leaves errors of about 1e-7 of the peak amplitude, |dlog10 I|
           about 1e-3 at voxels 1e-10 of the peak, and where lines are alike they repeat);
  * mixed: FP64 transforms, FP32 spectra between passes, estimate
           hypot(5.8e-3 / sqrt(N_last), 3.5e-3 / sqrt(N1 N_last)) (a term vanishes on one cell);
  * fp64:  FP64 transforms and spectra, the FP32 inputs widened on the device.
The default is mixed where its estimate is at most 6e-4, else fp64, with the pass order of the
smaller estimate among the orders that keep source groups; a mixed plan then gives FP32
transforms to pass 1, then pass 2, while the estimate with a worst-case FP32-transform term of
9.8e-3 sqrt(log2 N / N) for each stays within 6e-4 (strided passes of about 3200 points and more,
whose prime factors are at most 17).  Each arithmetic and order is
checked against the oracle: 1e-5 of the peak everywhere; |dlog10 I| <= 1e-3 for the ones the
plan would choose; fp64 at the reference's FP64 parity bound (rtol 1e-9, 1e-12 of the peak;
tests/test_mgpu.py:51-53 of the reference) against the oracle fed the same FP32-rounded order
parameter.
"""

import math

import numpy as np
import pytest

from helpers import check_fp32, dlog10, rel_to_peak
from qpop_diffraction import workloads
from qpop_diffraction.batch import PlanCache, Reflection, resident_job

pytestmark = pytest.mark.gpu


def estimate(dims, pd1):
    """The plan's error estimate of the mixed arithmetic with pass 1 along memory dim pd1."""
    nl, n1 = dims[3 - pd1], dims[pd1]
    spread = lambda k, n: k / math.sqrt(n) if n > 1 else 0.0
    return math.hypot(spread(5.8e-3, nl), spread(3.5e-3, n1 * nl))


def expected(dims):
    """(pass dims, bits of each pass's transforms, storage bits) of the default plan (PbTiO3:
    source groups in either order)."""
    pd = (0, 2, 1) if estimate(dims, 2) < estimate(dims, 1) else (0, 1, 2)
    est = estimate(dims, pd[1])
    if est > 6e-4:
        return pd, (64, 64, 64), 64
    bits = [64, 64, 64]
    def butterflies(n):  # every prime factor at most 17 (no direct stage)
        for q in (2, 3, 5, 7, 11, 13, 17):
            while n % q == 0:
                n //= q
        return n == 1

    for d in (1, 2):
        n = dims[pd[d]]
        tx = 9.8e-3 * math.sqrt(math.log2(n) / n) if n > 1 and butterflies(n) else 1.0
        if math.hypot(est, tx) <= 6e-4:
            bits[d] = 32
            est = math.hypot(est, tx)
    return pd, tuple(bits), 32


@pytest.mark.parametrize("dims", [(512, 512, 104), (128, 128, 30), (448, 448, 88), (256, 256, 64),
                                  (2000, 20, 1), (1, 20, 2000), (1024, 1024, 1), (68, 4000, 4000),
                                  (104, 512, 512), (40, 32, 20)])
def test_plan_choice(dims):
    refl = Reflection((0, 0, 2), workloads.pto_atoms((0, 0, 2)))
    snap = workloads.snapshot_spec(*dims)
    plan = PlanCache(32).get(snap, 0, refl)
    pd, pb, sb = expected(dims)
    print(f"{dims}: passes {plan.pass_dims} bits {plan.pass_bits}/{plan.storage_bits} "
          f"estimate {plan.dlog10_estimate:.1e}", flush=True)
    assert plan.pass_dims == pd
    assert (plan.pass_bits, plan.storage_bits) == (pb, sb)
    assert plan.compute_bits == pb[0]
    assert 0.0 <= plan.dlog10_estimate <= 6e-4


def run(snap, refl, order, arith):
    import torch
    plans = PlanCache(32, order=order, arith=arith)
    plan = plans.get(snap, 0, refl)
    out = torch.empty(plan.n, dtype=torch.float64, device="cuda")
    work = torch.empty(plan.workspace_bytes, dtype=torch.uint8, device="cuda")
    resident_job(plans, snap, 0, refl, out, work)
    nx, ny, nz = snap.dims
    return out.cpu().numpy().reshape(nz, ny, nx).transpose(2, 1, 0), plan


@pytest.mark.parametrize("order", ["memory", "swapped"])
@pytest.mark.parametrize("kind,shape,hkl", [("twin", (64, 64, 30), (1, 1, 3)), ("hex", (96, 80, 32), (1, 0, 3))])
def test_each_arith_against_oracle(kind, shape, hkl, order):
    """A thin film (short last axis in memory order) in every arithmetic and both pass orders."""
    from qpd_oracle import diffraction_fp64
    refl = Reflection(hkl, workloads.pto_atoms(hkl))
    make = workloads.twin_snapshot if kind == "twin" else workloads.hex_snapshot
    snap = make((kind,) + shape, *shape, seed=21)
    oPhase, P, U, region = workloads.host_inputs(snap)  # P holds the FP32 values the plan reads
    ref = diffraction_fp64(snap.params, refl.atoms, oPhase, P, U, region)[0]
    res = {}
    for arith in ("fp32", "mixed", "fp64"):
        I, plan = run(snap, refl, order, arith)
        res[arith] = (rel_to_peak(I, ref), dlog10(I, ref), plan.dlog10_estimate)
        print(f"{kind} {shape} {order} {arith}: bits {plan.compute_bits}/{plan.storage_bits} max|dI|/peak "
              f"{res[arith][0]:.2e} max|dlog10 I| {res[arith][1]:.2e} (estimate {res[arith][2]:.1e})", flush=True)
        assert res[arith][0] <= 1e-5
        assert np.argmax(I) == np.argmax(ref)
        if arith == "fp64":
            np.testing.assert_allclose(I, ref, rtol=1e-9, atol=1e-12 * ref.max())
        elif arith == "mixed" and res[arith][2] <= 6e-4:
            # the arithmetic the plan chooses on this grid meets the bound
            assert res[arith][1] <= 1e-3
    # FP64 transforms remove the FP32 transforms' errors
    assert res["mixed"][1] <= res["fp32"][1]
    I, plan = run(snap, refl, "auto", "auto")
    check_fp32(I, ref, f"default plan {plan.pass_dims} bits {plan.compute_bits}/{plan.storage_bits}")


@pytest.mark.parametrize("case", ["fraction_region", "one_hot"])
def test_wide_inputs_with_weights(case):
    """The fp64 arithmetic widens P, the phase fractions and the region weight: a two-phase
    volume with fractional weights and a region (the fraction path) and with one-hot phases
    and non-scattering cells, at the FP64 parity bound against the oracle."""
    import torch
    from qpd_oracle import diffraction_fp64
    from qpop_diffraction import Engine, Volume
    nx, ny, nz = 40, 24, 20
    w = workloads.two_phase(nx, ny, nz, n_grains=6, seed=4)
    n = nx * ny * nz
    gen = torch.Generator(device="cuda").manual_seed(12)
    phase, fraction, region = w.phase.clone(), None, None
    if case == "one_hot":
        phase[torch.rand(n, generator=gen, device="cuda") < 0.2] = -1
    else:
        fraction = torch.rand((2, n), generator=gen, device="cuda")
        fraction[fraction < 0.15] = 0.0
        region = torch.rand(n, generator=gen, device="cuda")
        phase = torch.zeros(n, dtype=torch.int8, device="cuda")
    eng = Engine(w.params, w.atoms, arith="fp64")
    vol = Volume(w.dims, w.axis, w.P, w.U, phase, fraction, region, 32)
    plan = eng.plan(vol.dims, vol.axis, 32)
    assert (plan.compute_bits, plan.storage_bits) == (64, 64)
    I = eng.intensity(vol).cpu().numpy().reshape(nz, ny, nx)
    to_xyz = lambda t, k: t.double().reshape(k, nz, ny, nx).permute(0, 3, 2, 1).contiguous().cpu().numpy()
    P = to_xyz(w.P, 3).ravel()
    U = to_xyz(w.U, 3).ravel()
    if fraction is None:
        ph = to_xyz(phase, 1)[0]
        oPhase = np.stack([(ph == p).astype(np.float64) for p in range(2)]).ravel()
        reg = (ph >= 0).astype(np.float64).ravel()
    else:
        oPhase = to_xyz(fraction, 2).ravel()
        reg = to_xyz(region, 1).ravel()
    ref = diffraction_fp64(w.params, w.atoms, oPhase, P, U, reg)[0].transpose(2, 1, 0)
    np.testing.assert_allclose(I, ref, rtol=1e-9, atol=1e-12 * ref.max())


def test_long_pass_fp32_transforms():
    """A 3300-point strided axis: the default plan gives that pass FP32 transforms (its
    worst-case term keeps the estimate within 6e-4) and the volume meets the FP32 bounds
    against the oracle; the same grid with FP64 transforms everywhere beside it."""
    from qpd_oracle import diffraction_fp64
    dims = (16, 3300, 40)
    pd, pb, sb = expected(dims)
    assert 32 in pb and pb[0] == 64, pb
    refl = Reflection((1, 0, 3), workloads.pto_atoms((1, 0, 3)))
    snap = workloads.twin_snapshot(("long",) + dims, *dims, seed=31)
    oPhase, P, U, region = workloads.host_inputs(snap)
    ref = diffraction_fp64(snap.params, refl.atoms, oPhase, P, U, region)[0]
    for arith in ("auto", "mixed"):
        I, plan = run(snap, refl, "auto", arith)
        print(f"{dims} {arith}: passes {plan.pass_dims} bits {plan.pass_bits} estimate "
              f"{plan.dlog10_estimate:.1e}", flush=True)
        if arith == "auto":
            assert plan.pass_bits == pb
        else:
            assert plan.pass_bits == (64, 64, 64)
        check_fp32(I, ref, f"{dims} {arith} bits {plan.pass_bits}")
