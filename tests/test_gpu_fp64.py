"""FP64 plans: source groups and every launch schedule against the FP64 oracle.

PbTiO3 has two source groups (Pb; Ti and the three O share su) and four y-classes, so FP64
plans, like FP32 ones, transform the groups in pass 0 and mix them in pass 1 (the
run-time-radix pass_strided_mix); QPD_CLASS_FORM=1 keeps the per-class form.  Every schedule
of every pass, in both forms, must meet the reference's parity bound between its two
implementations (rtol 1e-9 with an absolute floor of 1e-12 of the peak,
tests/test_mgpu.py:51-53 of the reference) on voxels of the full volume.
"""

import numpy as np
import pytest

from qpop_diffraction import workloads
from qpop_diffraction._native import fast_lengths
from qpop_diffraction.batch import PlanCache, Reflection, resident_job

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def refl():
    return Reflection((1, 0, 3), workloads.pto_atoms((1, 0, 3)))


def run_fp64(refl, shape, schedules=None, pick=None, order="memory"):
    """FP64 intensity of a twin snapshot; pick: {pass: suffix} selects the first schedule
    whose name ends with it; order: the plan's pass order (memory order: pass d along dim d)."""
    import torch
    nx, ny, nz = shape
    snap = workloads.twin_snapshot(("fp64",) + tuple(shape), nx, ny, nz, seed=5, precision=64)
    plans = PlanCache(64, order=order)
    plan = plans.get(snap, 0, refl)
    for d, idx in (schedules or {}).items():
        plan.set_schedule(d, idx)
    for d, suffix in (pick or {}).items():
        plan.set_schedule(d, next(i for i, nm in enumerate(plan.schedules(d)) if nm.endswith(suffix)))
    out = torch.empty(nx * ny * nz, dtype=torch.float64, device="cuda")
    work = torch.empty(plan.workspace_bytes, dtype=torch.uint8, device="cuda")
    resident_job(plans, snap, 0, refl, out, work)
    I = out.cpu().numpy().reshape(nz, ny, nx).transpose(2, 1, 0)
    return I, snap, plan


@pytest.mark.parametrize("order", ["auto", "memory"])
@pytest.mark.parametrize("shape", [(48, 40, 24), (104, 64, 30), (40, 24, 104)])
@pytest.mark.parametrize("form", ["groups", "class"])
def test_fp64_every_schedule(refl, shape, form, order, monkeypatch):
    from qpd_oracle import diffraction_fp64
    if form == "class":
        monkeypatch.setenv("QPD_CLASS_FORM", "1")
    else:
        monkeypatch.delenv("QPD_CLASS_FORM", raising=False)
    I0, snap, plan = run_fp64(refl, shape, order=order)
    ref = diffraction_fp64(snap.params, refl.atoms, *workloads.host_inputs(snap))[0]
    atol = 1e-12 * ref.max()
    np.testing.assert_allclose(I0, ref, rtol=1e-9, atol=atol)
    names = [plan.schedules(d) for d in range(3)]
    print(f"{shape} {form}: workspace {plan.workspace_bytes} B, schedules {names}", flush=True)
    for d in range(3):
        for idx in range(1, len(names[d])):
            I, _, _ = run_fp64(refl, shape, {d: idx}, order=order)
            np.testing.assert_allclose(I, ref, rtol=1e-9, atol=atol, err_msg=f"pass {d}: {names[d][idx]}")


def test_fp64_groups_use_less_workspace(refl, monkeypatch):
    """Source groups: pass 0 writes 3 fields per group (2 groups) instead of 3 per y-class (4)."""
    monkeypatch.delenv("QPD_CLASS_FORM", raising=False)
    _, _, grouped = run_fp64(refl, (48, 40, 24))
    monkeypatch.setenv("QPD_CLASS_FORM", "1")
    _, _, classes = run_fp64(refl, (48, 40, 24))
    assert grouped.workspace_bytes < classes.workspace_bytes


@pytest.mark.parametrize("shape", [(64, 48, 1), (1, 32, 20)])
def test_fp64_thin_volumes(refl, shape):
    """FP64 volumes one cell thick along an axis, at the reference's parity bound."""
    from qpd_oracle import diffraction_fp64
    I, snap, _ = run_fp64(refl, shape)
    ref = diffraction_fp64(snap.params, refl.atoms, *workloads.host_inputs(snap))[0]
    np.testing.assert_allclose(I, ref, rtol=1e-9, atol=1e-12 * ref.max())


@pytest.mark.parametrize("role,dim,form", [("source_wave", 0, "groups"), ("source", 0, "class"),
                                           ("strided", 1, "class"), ("groups", 1, "groups"), ("intensity", 2, "groups")])
def test_fp64_every_length(refl, role, dim, form, monkeypatch):
    """The compile-time FP64 kernels of a role on every length that has an FP32 kernel of the
    role, at the reference's bound.  A length whose FP64 tile does not fit (more than 256
    threads or 160 KB of LDS) runs the generic kernel, which must meet the same bound."""
    from qpd_oracle import diffraction_fp64
    if form == "class":
        monkeypatch.setenv("QPD_CLASS_FORM", "1")
    else:
        monkeypatch.delenv("QPD_CLASS_FORM", raising=False)
    lengths = fast_lengths("strided_groups2" if role == "groups" else role)
    ran_f64 = 0
    for N in lengths:
        # pass 1 of the source-group form needs a length with a source-group kernel on dim 1 (20)
        shape = [6, 20, 25] if role == "source_wave" else [6, 5, 4]
        shape[dim] = N
        plan_names = PlanCache(64, order="memory").get(
            workloads.twin_snapshot(("names",) + tuple(shape), *shape, seed=5, precision=64), 0, refl)
        names = plan_names.schedules(dim)
        wanted = " source-wave f64" if role == "source_wave" else " f64"
        pick = {dim: wanted} if any(nm.endswith(wanted) for nm in names) else None
        print(f"FP64 {role} N={N} grid {shape}: {'fast' if pick else 'generic'}", flush=True)
        I, snap, plan = run_fp64(refl, shape, pick=pick)
        ran_f64 += pick is not None
        ref = diffraction_fp64(snap.params, refl.atoms, *workloads.host_inputs(snap))[0]
        np.testing.assert_allclose(I, ref, rtol=1e-9, atol=1e-12 * ref.max(), err_msg=f"{role} N={N}")
    print(f"{role}: {ran_f64} of {len(lengths)} lengths on an FP64 kernel")
    assert ran_f64 >= len(lengths) // 2


@pytest.mark.parametrize("kind", ["source-wave", "source"])
@pytest.mark.parametrize("N,P", [(34, 8), (52, 3), (68, 4), (136, 2)])
def test_fp64_source_split_stage(refl, N, P, kind):
    """The FP64 source passes with a prime-radix stage split over P lanes per butterfly (the
    one-wave-per-row pass's last stage, the block-per-row pass's first), at the reference's
    bound."""
    from qpd_oracle import diffraction_fp64
    shape = (N, 20, 25)
    print(f"FP64 split {kind} N={N} P={P}", flush=True)
    I, snap, plan = run_fp64(refl, shape, pick={0: f" {kind} split{P} f64"})
    ref = diffraction_fp64(snap.params, refl.atoms, *workloads.host_inputs(snap))[0]
    np.testing.assert_allclose(I, ref, rtol=1e-9, atol=1e-12 * ref.max(), err_msg=f"split source N={N}")


@pytest.mark.parametrize("shape", [(6, 5, 4000), (6, 4000, 4), (6, 5, 4096), (4000, 3, 2)])
def test_fp64_long_axes(refl, shape):
    """Long FP64 axes: strided ones (FP64 plans do not split them) of 4000 (radix-5 stages, 800
    butterflies a line, which the all-radix kernels hold at 3 a thread) and 4096, the longest
    (radix 8 only: 512 butterflies a line, one a thread in the 8-values-a-thread kernels), and
    a contiguous 4000 (the source pass at eight cells a thread)."""
    from qpd_oracle import diffraction_fp64
    print(f"FP64 long axis {shape}", flush=True)
    I, snap, _ = run_fp64(refl, shape)
    ref = diffraction_fp64(snap.params, refl.atoms, *workloads.host_inputs(snap))[0]
    np.testing.assert_allclose(I, ref, rtol=1e-9, atol=1e-12 * ref.max())
