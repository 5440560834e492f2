"""The large-volume path at a size a full FP64 plan still handles.

* ``large.fp64_crop`` (x slabs through passes 0 and 1, the crop's ky kept, pass 2 on the
  reduced grid) equals the crop of a full FP64 plan of the same volume to 1e-12 of the peak;
* the polar-vortex generator's FP32 volume agrees with that FP64 crop to the FP32 bounds
  (1e-5 of the peak, |dlog10 I| <= 1e-3, same peak voxel);
* a long last axis runs in one launch: a split stores its first step in FP32, whose
  rounding its 25- to 128-point second step would spread over too few voxels (pass 2 of
  an FP32 plan of the automatic choice keeps no split schedule).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_fp64_crop_matches_full_plan():
    import torch
    from qpop_diffraction import Engine, Volume, workloads
    from qpop_diffraction.large import crop_of, fp64_crop
    v = workloads.vortex(240, 200, 20, substrate=4, slab=64, precision=64)
    print("vortex 240x200x20 fp64 crop vs full", flush=True)
    eng = Engine(v.params, v.atoms)
    vol = Volume(v.dims, v.axis, v.P, v.U, v.phase, precision=64)
    full = eng.intensity(vol)
    torch.cuda.synchronize()
    ref = crop_of(full, v.dims, crop=(64, 80))
    got = fp64_crop(v, v.params, eng.tables, crop=(64, 80), slab=50)
    peak = ref.max().item()
    err = (got - ref).abs().max().item() / peak
    print(f"  max|dI|/peak {err:.2e}")
    assert err < 1e-12
    assert int(got.argmax()) == int(ref.argmax())


def test_vortex_fp32_against_fp64_crop():
    import torch
    from qpop_diffraction import Engine, Volume, workloads
    from qpop_diffraction.large import crop_of, fp64_crop
    v = workloads.vortex(1000, 640, 68, precision=32)
    print("vortex 1000x640x68 fp32 vs fp64 crop", flush=True)
    eng = Engine(v.params, v.atoms)
    vol = Volume(v.dims, v.axis, v.P, v.U, v.phase, precision=32)
    plan = eng.plan(vol.dims, vol.axis, 32)
    names = plan.schedules(2)
    assert not any(nm.startswith("split") for nm in names), names
    I32 = crop_of(eng.intensity(vol), v.dims, crop=(128, 128))
    ref = fp64_crop(v, v.params, eng.tables, crop=(128, 128), slab=250)
    peak = ref.max()
    rel = ((I32 - ref).abs().max() / peak).item()
    mask = ref >= 1e-10 * peak
    dl = (torch.log10(I32[mask]) - torch.log10(ref[mask])).abs().max().item()
    print(f"  max|dI|/peak {rel:.2e}, max|dlog10 I| {dl:.2e}")
    assert rel <= 1e-5
    assert dl <= 1e-3
    assert int(I32.argmax()) == int(ref.argmax())


def test_small_vortex_on_substrate_against_oracle():
    """The polar-vortex generator with its 16-cell non-scattering substrate (z-fastest layout,
    u_z ramp of the substrate's misfit) at a size the oracle handles, FP32 default plan against
    the FP64 oracle at the FP32 bounds, and the FP64 plan at the reference's bound."""
    import torch
    from helpers import check_fp32
    from qpd_oracle import diffraction_fp64
    from qpop_diffraction import Engine, Volume, workloads
    nx, ny, nz = 80, 64, 40
    print(f"vortex {nx}x{ny}x{nz} on a 16-cell substrate vs oracle", flush=True)
    out = {}
    for prec in (32, 64):
        v = workloads.vortex(nx, ny, nz, substrate=16, slab=40, precision=prec)
        eng = Engine(v.params, v.atoms)
        vol = Volume(v.dims, v.axis, v.P, v.U, v.phase, precision=prec)
        out[prec] = eng.intensity(vol).cpu().numpy().reshape(nx, ny, nz)
    # the oracle's inputs: (x, y, z) C order is the generator's z-fastest layout itself
    n = nx * ny * nz
    P = v.P.double().cpu().numpy().ravel()
    U = v.U.cpu().numpy().ravel()
    ph = v.phase.cpu().numpy()
    oPhase = (ph == 0).astype(np.float64)
    region = (ph >= 0).astype(np.float64)
    v32 = workloads.vortex(nx, ny, nz, substrate=16, slab=40, precision=32)
    P32 = v32.P.double().cpu().numpy().ravel()
    ref = diffraction_fp64(v.params, v.atoms, oPhase, P, U, region)[0]
    ref32 = diffraction_fp64(v.params, v.atoms, oPhase, P32, U, region)[0]
    np.testing.assert_allclose(out[64], ref, rtol=1e-9, atol=1e-12 * ref.max())
    check_fp32(out[32], ref32, "vortex on substrate")
    assert n == out[32].size


def test_full_size_vortex_fp32_against_fp64_crop():
    """The 4000 x 4000 x 68 polar-vortex volume whole on one GPU with the default FP32 plan,
    its 512 x 512 x 68 crop around G against the FP64 crop at the FP32 bounds."""
    import torch
    from qpop_diffraction import Engine, Volume, workloads
    from qpop_diffraction.large import crop_of, fp64_crop
    torch.cuda.empty_cache()
    free, total = torch.cuda.mem_get_info()
    if free < 180e9:
        pytest.skip(f"needs about 140 GB of free HBM, {free / 1e9:.0f} GB free")
    print("vortex 4000x4000x68 fp32 vs fp64 crop", flush=True)
    v = workloads.vortex(4000, 4000, 68, precision=32)
    eng = Engine(v.params, v.atoms)
    vol = Volume(v.dims, v.axis, v.P, v.U, v.phase, precision=32)
    I = eng.intensity(vol)
    I32 = crop_of(I, v.dims, crop=(512, 512))
    del I
    eng._work.clear()
    torch.cuda.empty_cache()
    ref = fp64_crop(v, v.params, eng.tables, crop=(512, 512), slab=250)
    del v, vol
    torch.cuda.empty_cache()
    peak = ref.max()
    rel = ((I32 - ref).abs().max() / peak).item()
    mask = ref >= 1e-10 * peak
    dl = (torch.log10(I32[mask]) - torch.log10(ref[mask])).abs().max().item()
    print(f"  max|dI|/peak {rel:.2e}, max|dlog10 I| {dl:.2e} over {int(mask.sum())} voxels", flush=True)
    assert rel <= 1e-5
    assert dl <= 1e-3
    assert int(I32.argmax()) == int(ref.argmax())
    del I32, ref
    torch.cuda.empty_cache()
