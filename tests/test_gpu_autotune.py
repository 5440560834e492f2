"""The schedule autotuner: it picks one of the plan's schedules for every pass, the
tuned volume stays within the FP32 bound of the oracle (1e-5 of the peak, same peak
voxel), and a cached choice is applied again without timing and gives the same bits."""

import numpy as np
import pytest

from helpers import check_fp32
from qpop_diffraction import workloads
from qpop_diffraction.autotune import ScheduleCache, tune
from qpop_diffraction.batch import Batch, PlanCache, Reflection

pytestmark = pytest.mark.gpu


def test_tune_and_cache(tmp_path):
    import torch
    from qpd_oracle import diffraction_fp64
    refl = Reflection((1, 1, 3), workloads.pto_atoms((1, 1, 3)))
    snap = workloads.twin_snapshot(("tune",), 48, 1000, 20, seed=4)
    plan = PlanCache(32).get(snap, 0, refl)
    n = 48 * 1000 * 20
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    work = torch.empty(plan.workspace_bytes, dtype=torch.uint8, device="cuda")
    cache = ScheduleCache(tmp_path / "schedules.json")
    rep = tune(plan, snap.P, snap.U, snap.phase, out, work, cache=cache)
    print(rep["passes"])
    assert not rep["cached"]
    for d in range(3):
        assert plan.schedules(d)[plan.schedule(d)] == rep["passes"][d]["chosen"]
    torch.cuda.synchronize()
    I1 = out.clone()
    I = I1.cpu().numpy().reshape(20, 1000, 48).transpose(2, 1, 0)
    ref = diffraction_fp64(snap.params, refl.atoms, *workloads.host_inputs(snap))[0]
    check_fp32(I, ref, "tuned plan")
    # a new plan of the same shape takes the stored choice
    plan2 = PlanCache(32).get(snap, 0, refl)
    out2 = torch.empty_like(out)
    rep2 = tune(plan2, snap.P, snap.U, snap.phase, out2, work, cache=ScheduleCache(tmp_path / "schedules.json"))
    assert rep2["cached"]
    assert [plan2.schedule(d) for d in range(3)] == [plan.schedule(d) for d in range(3)]
    torch.cuda.synchronize()
    assert torch.equal(out2, I1)


def test_tuned_batch_matches_untuned_bound():
    import torch
    refls = [Reflection(h, workloads.pto_atoms(h)) for h in workloads.REFLECTIONS[:2]]
    snaps = [workloads.twin_snapshot(("tb", i), *g, seed=20 + i) for i, g in enumerate([(64, 48, 30), (40, 40, 68)])]
    b = Batch(snaps, refls, n_streams=2, graph=True, tune=True, tune_cache=ScheduleCache(None))
    assert len(b.tuning) == len(b.plans)
    b.capture()
    b.run()
    torch.cuda.synchronize()
    ref = Batch(snaps, refls, n_streams=1, graph=False)
    ref.capture()
    torch.cuda.synchronize()
    for j, k in zip(b.jobs, ref.jobs):
        peak = k.out.max()
        assert ((j.out - k.out).abs().max() / peak).item() <= 2e-6
        assert int(j.out.argmax()) == int(k.out.argmax())
