"""Every compile-time-length kernel and every launch schedule against numpy.fft / the oracle.

* qpd_fft3d on each registered length of the plain contiguous and strided roles
  (FP32, to 2e-6 of the largest output, like test_gpu_fft.py);
* the volume path with each registered length on each memory dim it has a kernel for
  (source pass on dim 0, 3 -> 2 pass on dim 1, intensity pass on dim 2), with the passes
  in memory order and every pass in FP32 so that the FP32 kernel of that length runs,
  against the FP64 oracle: max |dI| <= 1e-5 of the peak and the same peak voxel; the same
  with passes 1 and 2 in FP64 (the pass-1 kernels that read FP32 spectra);
* long strided lengths with no kernel of their own run split into two launches
  (N = N1 x N2 up to 4000, the length of the 4000 x 4000 x 68 volume), same bound;
* on one plan, every schedule of every pass gives the same bound;
* every volume above also runs through the default plan (pass order and arithmetic chosen
  by the plan), held to the FP32 bounds: 1e-5 of the peak, |dlog10 I| <= 1e-3, same peak.
"""

import numpy as np
import pytest

from helpers import FP32_PEAK_TOL, check_fp32, rel_to_peak
from qpop_diffraction import workloads
from qpop_diffraction._native import fast_lengths
from qpop_diffraction.batch import PlanCache, Reflection, resident_job

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def refl():
    return Reflection((1, 0, 3), workloads.pto_atoms((1, 0, 3)))


def fft_check(shape, nf=2):
    import torch
    from qpop_diffraction._native import fft3d
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal((nf,) + shape) + 1j * rng.standard_normal((nf,) + shape)
    t = torch.from_numpy(x).to(dtype=torch.complex64, device="cuda").contiguous()
    fft3d(t)
    got = t.cpu().numpy()
    ref = np.fft.fftn(x, axes=(1, 2, 3))
    return np.max(np.abs(got - ref)) / np.max(np.abs(ref))


def test_fft3d_every_contiguous_length():
    worst = 0.0
    for N in fast_lengths("contiguous"):
        err = fft_check((2, 3, N))
        worst = max(worst, err)
        assert err < 2e-6, (N, err)
    print(f"contiguous lengths: worst {worst:.2e}")


def test_fft3d_every_strided_length():
    worst = 0.0
    for N in fast_lengths("strided1"):
        for shape in ((N, 3, 5), (3, N, 5)):
            err = fft_check(shape, nf=1)
            worst = max(worst, err)
            assert err < 2e-6, (shape, err)
    print(f"strided lengths: worst {worst:.2e}")


def volume_check(refl, nx, ny, nz, seed=0, expect=None, schedules=None, pick=None, order="memory", arith="fp32",
                 strict=False):
    """FP32 volume of a twin snapshot against the oracle, with the kernels under test selected;
    returns max |dI| / peak.  The same snapshot through the default plan is held to the FP32
    bounds.

    order, arith: plan options (by default the memory order and FP32 in every pass, so that
    pass d runs along memory dim d in FP32); expect: {pass: prefix of the default schedule's
    name}; pick: {pass: suffix} selects the first schedule whose name ends with it (which must
    exist); strict: the plan under test is also held to the FP32 bounds (|dlog10 I| included)."""
    import torch
    from qpd_oracle import diffraction_fp64
    snap = workloads.twin_snapshot(("len", nx, ny, nz), nx, ny, nz, seed=seed)
    plans = PlanCache(32, order=order, arith=arith)
    plan = plans.get(snap, 0, refl)
    if expect is not None:
        for d, prefix in expect.items():
            names = plan.schedules(d)
            assert names[plan.schedule(d)].startswith(prefix), (d, names)
    if pick is not None:
        for d, suffix in pick.items():
            names = plan.schedules(d)
            idx = next(i for i, nm in enumerate(names) if nm.endswith(suffix))
            plan.set_schedule(d, idx)
    if schedules is not None:
        for d, idx in schedules.items():
            plan.set_schedule(d, idx)
    out = torch.empty(nx * ny * nz, dtype=torch.float64, device="cuda")
    work = torch.empty(plan.workspace_bytes, dtype=torch.uint8, device="cuda")
    resident_job(plans, snap, 0, refl, out, work)
    I = out.cpu().numpy().reshape(nz, ny, nx).transpose(2, 1, 0)
    ref = diffraction_fp64(snap.params, refl.atoms, *workloads.host_inputs(snap))[0]
    assert np.argmax(I) == np.argmax(ref), (nx, ny, nz)
    if strict:
        check_fp32(I, ref, f"plan under test {(nx, ny, nz)} {[plan.schedules(d)[plan.schedule(d)] for d in range(3)]}")
    dplans = PlanCache(32)
    dplan = dplans.get(snap, 0, refl)
    dwork = torch.empty(dplan.workspace_bytes, dtype=torch.uint8, device="cuda")
    resident_job(dplans, snap, 0, refl, out, dwork)
    Id = out.cpu().numpy().reshape(nz, ny, nx).transpose(2, 1, 0)
    check_fp32(Id, ref, f"default plan {(nx, ny, nz)} passes {dplan.pass_dims} "
                        f"bits {dplan.compute_bits}/{dplan.storage_bits}")
    return rel_to_peak(I, ref)


@pytest.mark.parametrize("role,dim,form,arith", [("source_wave", 0, "groups", "fp32"), ("source", 0, "groups", "fp32"),
                                                 ("source", 0, "class", "fp32"), ("strided", 1, "class", "fp32"),
                                                 ("groups", 1, "groups", "fp32"), ("intensity", 2, "groups", "fp32"),
                                                 ("source_wave", 0, "groups", "mixed"),
                                                 ("source", 0, "class", "mixed"),
                                                 ("strided", 1, "class", "mixed"),
                                                 ("groups", 1, "groups", "mixed"),
                                                 ("intensity", 2, "groups", "mixed"),
                                                 ("strided", 1, "class", "mixed-1buf"),
                                                 ("intensity", 2, "groups", "mixed-1buf")])
def test_volume_every_length(refl, role, dim, form, arith, monkeypatch):
    """Pass 1 runs the source-group kernel for PbTiO3 unless QPD_CLASS_FORM=1; pass 0 of
    the source-group form defaults to the one-wave-per-row kernel where the length has one,
    and the block-per-row source kernel is picked explicitly for the "source" case.  With
    arith "mixed" the pass runs the FP64 kernel of the length on FP32 arrays (FP32 inputs and
    spectra), "mixed-1buf" its variant with one LDS stage buffer; a length whose FP64 tile
    does not fit runs the generic kernel."""
    if form == "class":
        monkeypatch.setenv("QPD_CLASS_FORM", "1")
    else:
        monkeypatch.delenv("QPD_CLASS_FORM", raising=False)
    worst = 0.0
    lengths = fast_lengths("strided_groups2" if role == "groups" else role)
    suffix = None
    if arith == "mixed-1buf":
        arith, suffix = "mixed", " f64 io32 1buf"
    elif arith == "mixed":
        suffix = " source-wave f64 io32" if role == "source_wave" else " f64 io32"
    ran_fast = 0
    for N in lengths:
        # pass 1 of the source-group form needs a compile-time length on dim 1 (20); 500 rows
        # leave a partial last block for most rows-per-block of the source kernels
        shape = [6, 20, 25] if role.startswith("source") and form == "groups" else [6, 5, 4]
        shape[dim] = N
        print(f"{role} N={N} grid {shape}", flush=True)
        if suffix:
            # a length whose FP64 tile does not fit (more than 256 threads) runs the generic kernel
            names = PlanCache(32, order="memory", arith=arith).get(
                workloads.twin_snapshot(("names",) + tuple(shape), *shape, seed=0), 0, refl).schedules(dim)
            has = any(nm.endswith(suffix) for nm in names)
            ran_fast += has
            err = volume_check(refl, *shape, pick={dim: suffix} if has else None, arith=arith)
        elif role == "source_wave":
            err = volume_check(refl, *shape, expect={dim: "fast"}, pick={dim: " source-wave"})
        elif role == "source" and form == "groups":
            err = volume_check(refl, *shape, expect={dim: "fast"}, pick={dim: " source"})
        else:
            err = volume_check(refl, *shape, expect={dim: "fast"})
        worst = max(worst, err)
        assert err <= FP32_PEAK_TOL, (N, err)
    print(f"{role} {arith}: worst max|dI|/peak {worst:.2e}")
    if suffix:
        print(f"{role}: {ran_fast} of {len(lengths)} lengths on an FP64 kernel reading FP32 spectra")
        assert ran_fast >= len(lengths) // 2


@pytest.mark.parametrize("arith", ["mixed", "fp32"])
def test_held_staged_every_length(refl, arith):
    """The held pass 1 with its next input staged in LDS (sixteen-byte loads of two columns of
    a 16-column tile; FP64 transforms on FP32 spectra, and FP32 throughout) on every pass-1
    length that has it, against the oracle (FP64 transforms: the FP32 bounds including
    |dlog10 I|; FP32: 1e-5 of the peak): 32 columns (two whole tiles) and 22 (a partial second
    tile, whose columns beyond the grid are neither loaded nor stored).  Grids with an odd
    number of columns do not list it."""
    ran = 0
    for N in fast_lengths("strided_groups2"):
        names = PlanCache(32, order="memory", arith=arith).get(
            workloads.twin_snapshot(("names", 32, N, 4), 32, N, 4, seed=0), 0, refl).schedules(1)
        if not any(nm.endswith(" ldsin") for nm in names):
            continue
        for nx in (32, 22):
            print(f"held staged N={N} {arith} nx={nx}", flush=True)
            err = volume_check(refl, nx, N, 4, pick={1: " ldsin"}, arith=arith, strict=arith == "mixed")
            assert err <= FP32_PEAK_TOL, (N, nx, err)
        ran += 1
        odd = (23, N, 4)
        onames = PlanCache(32, order="memory", arith=arith).get(
            workloads.twin_snapshot(("names",) + odd, *odd, seed=0), 0, refl).schedules(1)
        assert not any(nm.endswith(" ldsin") for nm in onames), (N, onames)
    print(f"{ran} lengths with a staged held pass")
    assert ran >= 4


# lengths whose source passes also have a prime-radix stage shared by several lanes per
# butterfly, with that lane count: 34 = 2 x 17 (eight lanes), 52 = 4 x 13 (three), 68 = 4 x 17
# (four), 136 = 8 x 17 (two); the one-wave-per-row pass splits its last stage (radix order
# 2x17, 4x13, ...), the block-per-row pass its first (17x2, 13x4, ...)
SPLIT_SOURCE = [(34, 8), (52, 3), (68, 4), (136, 2)]


@pytest.mark.parametrize("kind", ["source-wave", "source"])
@pytest.mark.parametrize("arith", ["fp32", "mixed"])
@pytest.mark.parametrize("N,P", SPLIT_SOURCE)
def test_source_split_stage(refl, N, P, arith, kind):
    """The source passes with a stage split over P lanes per butterfly (each lane forms the
    outputs of one coset of the radix's multiplicative group) against the oracle, FP32 and
    FP64 transforms on FP32 arrays, on rows that leave a partial last block.  The FP64
    transforms are also held to |dlog10 I| <= 1e-3 (explicit FP32 arithmetic has no such
    bound)."""
    suffix = f" {kind} split{P}" + (" f64 io32" if arith == "mixed" else "")
    shape = [N, 20, 25]
    print(f"split {kind} N={N} P={P} {arith}", flush=True)
    err = volume_check(refl, *shape, pick={0: suffix}, arith=arith, strict=arith == "mixed")
    assert err <= FP32_PEAK_TOL, (N, err)


@pytest.mark.parametrize("N", [640, 1000, 2000, 4000])
@pytest.mark.parametrize("dim", [1, 2])
def test_split_long_strided(refl, N, dim):
    shape = [5, 4, 3]
    shape[dim] = N
    print(f"split N={N} on dim {dim}, grid {shape}", flush=True)
    err = volume_check(refl, *shape, expect={dim: "split"})
    print(f"  max|dI|/peak {err:.2e}")
    assert err <= FP32_PEAK_TOL


def test_intensity_staged_every_length(refl):
    """The FP64 intensity pass on FP32 spectra with its next input staged in LDS (an autotuner
    alternative) on every length that lists it, on 6 columns (one partial tile), against the
    oracle at the FP32 bounds including |dlog10 I|."""
    ran = 0
    for N in fast_lengths("intensity"):
        shape = (6, 5, N)
        names = PlanCache(32, order="memory", arith="mixed").get(
            workloads.twin_snapshot(("names",) + shape, *shape, seed=0), 0, refl).schedules(2)
        if not any(nm.endswith(" ldsin") for nm in names):
            continue
        print(f"intensity staged N={N}", flush=True)
        err = volume_check(refl, *shape, pick={2: " ldsin"}, arith="mixed", strict=True)
        assert err <= FP32_PEAK_TOL, (N, err)
        ran += 1
    print(f"{ran} lengths with a staged intensity pass")
    assert ran >= 4


@pytest.mark.parametrize("arith,dim,N", [("fp32", 1, 1000), ("fp32", 1, 4000), ("fp32", 2, 4000),
                                         ("mixed", 2, 2000), ("mixed", 2, 4000)])
@pytest.mark.parametrize("nx", [6, 18])
def test_split_staged(refl, arith, dim, N, nx):
    """Split long strided passes whose first step stages its next field in LDS (an even number
    of columns: 6 is one partial tile, 18 a whole one and a partial one for the FP64 step's
    eight columns), in FP32 and as the FP64 intensity split of FP64 transforms on FP32 spectra,
    against the oracle (FP64 transforms also at |dlog10 I| <= 1e-3)."""
    shape = [nx, 4, 3]
    shape[dim] = N
    print(f"staged split N={N} dim {dim} {arith} grid {shape}", flush=True)
    err = volume_check(refl, *shape, expect={dim: "split"}, arith=arith, strict=arith == "mixed")
    snap = workloads.twin_snapshot(("len",) + tuple(shape), *shape, seed=0)
    plan = PlanCache(32, order="memory", arith=arith).get(snap, 0, refl)
    name = plan.schedules(dim)[plan.schedule(dim)]
    print(f"  {name}: max|dI|/peak {err:.2e}")
    assert "ldsin" in name, name
    assert err <= FP32_PEAK_TOL


@pytest.mark.parametrize("N", [2000, 4000])
def test_split_intensity_fp64(refl, N):
    """The intensity pass of a plan computing in FP64 on FP32 spectra, split in two launches
    with an FP64 intermediate in the workspace (its first step reads the FP32 spectra)."""
    shape = [5, 4, N]
    print(f"fp64 split N={N} on dim 2, grid {shape}", flush=True)
    err = volume_check(refl, *shape, expect={2: "split"}, arith="mixed", strict=True)
    print(f"  max|dI|/peak {err:.2e}")
    assert err <= FP32_PEAK_TOL


@pytest.mark.parametrize("N", [3000, 4000])
def test_long_contiguous_rows(refl, N):
    """Rows longer than the compile-time lengths and than 2048 cells (the generic source pass
    at eight cells a thread), as in a 4000 x 4000 x 68 volume laid out x fastest by the CLI."""
    print(f"contiguous N={N}", flush=True)
    err = volume_check(refl, N, 3, 2, expect={0: "generic"})
    print(f"  max|dI|/peak {err:.2e}")
    assert err <= FP32_PEAK_TOL


@pytest.mark.parametrize("shape", [(40, 1000, 6), (24, 20, 500), (64, 20, 104)])
def test_every_schedule(refl, shape):
    import torch
    snap = workloads.twin_snapshot(("sched",) + shape, *shape, seed=1)
    plan = PlanCache(32, order="memory", arith="fp32").get(snap, 0, refl)
    counts = [len(plan.schedules(d)) for d in range(3)]
    print(f"{shape}: schedules per pass {counts}", flush=True)
    for d in range(3):
        for idx, name in enumerate(plan.schedules(d)):
            print(f"  pass {d} schedule {idx}: {name}", flush=True)
            err = volume_check(refl, *shape, seed=1, schedules={d: idx})
            assert err <= FP32_PEAK_TOL, (d, name, err)


@pytest.mark.parametrize("arith", ["mixed", "fp32"])
def test_every_schedule_held_pass(refl, arith):
    """Every schedule of every pass of a plan with pass 1 along 104 points (the held pass over
    two source groups: eight and sixteen columns a tile, two LDS stage buffers and one; 48
    columns leave a partial last tile), in FP64 transforms on FP32 arrays (held to the
    FP32 bounds including |dlog10 I|) and in FP32 transforms (1e-5 of the peak)."""
    shape = (48, 104, 12)
    snap = workloads.twin_snapshot(("len",) + shape, *shape, seed=1)
    plan = PlanCache(32, order="memory", arith=arith).get(snap, 0, refl)
    names = [plan.schedules(d) for d in range(3)]
    if arith == "mixed":
        assert any("G16 groups held2 f64 io32 1buf" in nm for nm in names[1]), names[1]
    for d in range(3):
        for idx, name in enumerate(names[d]):
            print(f"  pass {d} schedule {idx}: {name}", flush=True)
            err = volume_check(refl, *shape, seed=1, schedules={d: idx}, arith=arith, strict=arith == "mixed")
            assert err <= FP32_PEAK_TOL, (d, name, err)


@pytest.mark.parametrize("shape,axis", [((48, 40, 24), (0, 1, 2)), ((24, 40, 48), (2, 1, 0)), ((104, 64, 30), (0, 1, 2)),
                                        ((24, 1000, 12), (0, 1, 2))])
def test_source_groups_match_class_form(refl, shape, axis, monkeypatch):
    """PbTiO3 has two source groups (Pb; Ti and the three O share su) and four y-classes, so
    FP32 plans transform the groups in pass 0 and mix them in pass 1; QPD_CLASS_FORM=1 keeps
    the per-class form.  Both agree to FP32 rounding (2e-6 of the peak) and with the oracle."""
    import torch
    from qpd_oracle import diffraction_fp64
    snap = workloads.twin_snapshot(("grp",) + shape, *shape, seed=9)
    if axis != (0, 1, 2):
        nx, ny, nz = shape
        snap.P = snap.P.reshape(3, nz, ny, nx).permute(0, 3, 2, 1).contiguous().reshape(3, -1)
        snap.U = snap.U.reshape(3, nz, ny, nx).permute(0, 3, 2, 1).contiguous().reshape(3, -1)
        snap.phase = snap.phase.reshape(nz, ny, nx).permute(2, 1, 0).contiguous().reshape(-1)
        snap.dims, snap.axis = (nz, ny, nx), (2, 1, 0)
    outs = {}
    for form in ("groups", "class"):
        if form == "class":
            monkeypatch.setenv("QPD_CLASS_FORM", "1")
        else:
            monkeypatch.delenv("QPD_CLASS_FORM", raising=False)
        plans = PlanCache(32)
        plan = plans.get(snap, 0, refl)
        names = plan.schedules(1)
        assert ("groups" in names[plan.schedule(1)]) == (form == "groups"), names
        out = torch.empty(plan.n, dtype=torch.float64, device="cuda")
        work = torch.empty(plan.workspace_bytes, dtype=torch.uint8, device="cuda")
        resident_job(plans, snap, 0, refl, out, work)
        outs[form] = out.cpu().numpy()
    monkeypatch.delenv("QPD_CLASS_FORM", raising=False)
    rel = np.max(np.abs(outs["groups"] - outs["class"])) / np.max(outs["class"])
    print(f"{shape} axis {axis}: groups vs class max|dI|/peak {rel:.2e}")
    assert rel <= 2e-6
    nx, ny, nz = shape
    I = outs["groups"].reshape(nz, ny, nx) if axis == (0, 1, 2) else outs["groups"].reshape(nx, ny, nz).transpose(2, 1, 0)
    snap0 = workloads.twin_snapshot(("grp",) + shape, *shape, seed=9)
    ref = diffraction_fp64(snap0.params, refl.atoms, *workloads.host_inputs(snap0))[0].transpose(2, 1, 0)
    check_fp32(I, ref, f"{shape} axis {axis}")


@pytest.mark.parametrize("case", ["two_phase", "fraction_region"])
def test_source_schedules_with_weights(case):
    """Every pass-0 schedule of a source-group plan (the one-wave-per-row kernel, the
    block-per-row kernel, the generic one) on a two-phase volume: with non-scattering cells
    (one-hot phase index, -1 in places), and with fractional phase weights and a region
    weight (the fraction path).  Each against the oracle at the FP32 bounds."""
    import torch
    from qpd_oracle import diffraction_fp64
    from qpop_diffraction import Engine, Volume
    nx, ny, nz = 64, 24, 20
    w = workloads.two_phase(nx, ny, nz, n_grains=6, seed=4)
    n = nx * ny * nz
    gen = torch.Generator(device="cuda").manual_seed(11)
    phase, fraction, region = w.phase.clone(), None, None
    if case == "two_phase":
        phase[torch.rand(n, generator=gen, device="cuda") < 0.2] = -1
    else:
        fraction = torch.rand((2, n), generator=gen, device="cuda", dtype=torch.float64)
        fraction[fraction < 0.15] = 0.0
        region = torch.rand(n, generator=gen, device="cuda", dtype=torch.float64)
        phase = torch.zeros(n, dtype=torch.int8, device="cuda")
    eng = Engine(w.params, w.atoms)
    vol = Volume(w.dims, w.axis, w.P, w.U, phase, fraction, region, 32)
    plan = eng.plan(vol.dims, vol.axis, 32)
    names = plan.schedules(0)
    assert "source-wave" in names[plan.schedule(0)], names
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
    for idx, name in enumerate(names):
        plan.set_schedule(0, idx)
        I = eng.intensity(vol).cpu().numpy().reshape(nz, ny, nx)
        check_fp32(I, ref, f"{case} pass 0 schedule {name}")


@pytest.mark.parametrize("shape", [(64, 48, 1), (40, 1, 24), (1, 32, 20), (1, 1, 16)])
def test_thin_volumes(refl, shape):
    """Volumes one cell thick along an axis (the reference's 2-D runs, e.g. its 400x400x1
    polar-vortex example): a length-1 transform is the identity, FP32 against the oracle."""
    err = volume_check(refl, *shape, seed=3)
    print(f"{shape}: {err:.2e}", flush=True)
    assert err <= FP32_PEAK_TOL, (shape, err)
