"""qpd_intensity_noalloc (no workspace argument, static per-GPU scratch) against qpd_intensity.

Both entry points run the same kernels on the same inputs, so the intensities must be
bitwise equal.  Two calls in a row, the second on a smaller grid (the scratch keeps the first
call's contents), in FP32 (default arithmetic) and FP64 plans; a plan whose workspace exceeds
the scratch is rejected by the C ABI before anything is enqueued.
"""

import ctypes

import pytest

from qpop_diffraction import _native, workloads
from qpop_diffraction.batch import PlanCache, Reflection

pytestmark = pytest.mark.gpu


def header_scratch_bytes():
    import re
    text = (_native.LIB_PATH.parents[3] / "include" / "qpd.h").read_text()
    m = re.search(r"#define QPD_NOALLOC_SCRATCH_BYTES \(\(size_t\)(\d+) << (\d+)\)", text)
    return int(m.group(1)) << int(m.group(2))


def pair(plans, snap, refl):
    """(intensity with a workspace, intensity from the static scratch) of one job."""
    import torch
    plan = plans.get(snap, 0, refl)
    assert plan.workspace_bytes <= _native.NOALLOC_SCRATCH_BYTES
    ref = torch.empty(plan.n, dtype=torch.float64, device="cuda")
    got = torch.full((plan.n,), -1.0, dtype=torch.float64, device="cuda")
    work = torch.empty(plan.workspace_bytes, dtype=torch.uint8, device="cuda")
    plan.intensity(snap.P, snap.U, snap.phase, ref, work)
    plan.intensity_noalloc(snap.P, snap.U, snap.phase, got)
    torch.cuda.synchronize()
    return ref, got, plan


@pytest.mark.parametrize("precision", [32, 64])
def test_noalloc_matches_workspace_entry_point(precision):
    import torch
    assert header_scratch_bytes() == _native.NOALLOC_SCRATCH_BYTES
    refl = Reflection((1, 0, 3), workloads.pto_atoms((1, 0, 3)))
    plans = PlanCache(precision)
    # the second grid is smaller than the first: the scratch holds stale spectra of the first
    for shape, seed in (((64, 64, 32), 3), ((40, 24, 20), 4)):
        snap = workloads.hex_snapshot(("noalloc", precision) + shape, *shape, seed=seed, precision=precision)
        ref, got, plan = pair(plans, snap, refl)
        print(f"noalloc precision {precision} grid {shape}: workspace {plan.workspace_bytes} bytes, "
              f"bits {plan.compute_bits}/{plan.storage_bits}", flush=True)
        assert torch.isfinite(ref).all()
        assert torch.equal(ref, got)


def test_noalloc_rejects_plan_beyond_scratch():
    """A 512 x 512 x 104 plan needs about 1.3 GB of workspace: the C ABI refuses it with
    QPD_ERR_UNSUPPORTED before any launch (the buffers it is given are never read)."""
    import torch
    refl = Reflection((0, 0, 2), workloads.pto_atoms((0, 0, 2)))
    plan = PlanCache(32).get(workloads.snapshot_spec(512, 512, 104), 0, refl)
    assert plan.workspace_bytes > _native.NOALLOC_SCRATCH_BYTES
    small = torch.zeros(16, dtype=torch.float64, device="cuda")
    fields = _native.Fields(small.data_ptr(), small.data_ptr(), small.data_ptr(), None, None)
    lib = _native.load_library()
    rc = lib.qpd_intensity_noalloc(plan._h, ctypes.byref(fields), small.data_ptr(), _native.current_stream())
    assert rc == -2
    assert b"qpd_intensity_noalloc holds at most" in lib.qpd_last_error()
    with pytest.raises(ValueError):
        n = plan.n
        P = torch.zeros(3 * n, dtype=torch.float32, device="cuda")
        U = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
        ph = torch.zeros(n, dtype=torch.int8, device="cuda")
        plan.intensity_noalloc(P, U, ph, torch.empty(n, dtype=torch.float64, device="cuda"))
