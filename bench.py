#!/usr/bin/env python3
"""Benchmark of the kinematic-diffraction path on one MI355X.

Workload: a 512 x 512 x 104 two-phase volume (40 periodic Voronoi grains,
tetragonal P = +-0.6 z with c/a = 1.06 through the u_z ramp, or rhombohedral
P = 0.4 (+-1, +-1, +-1)/sqrt 3; u = 0.5 % in-plane ramp + Gaussian random field of
rms 0.005 nm), PbTiO3-like 5-atom cell, reflection 002, synthetic seeded data
generated on the GPU.  One step = one volume: source, 3-D transforms and the
fftshifted intensity (three kernel launches), inputs resident in HBM.

Prints one JSON line: voxels/s (FP32 path) as ``value``, ms per volume, the
FP64 path's ms per volume, |dlog10 I| of FP32 against FP64, per-pass kernel
times and GB/s, the roofline entry of the dominant kernel, and the CPU
baseline (the FP64 NumPy oracle on a 256 x 256 x 104 crop, one core).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "q-pop-diffraction_amd"))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, nargs=3, default=(512, 512, 104), metavar=("NX", "NY", "NZ"))
    ap.add_argument("--cpu-crop", type=int, default=256, help="lateral size of the CPU-baseline crop (0: skip)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_steps(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def pass_times(plan, vol, out, work, steps):
    """Average duration of each of the three kernels, HIP events on the stream they run on."""
    import torch
    stream = torch.cuda.current_stream()
    res = []
    for p in range(3):
        plan.run_pass(p, vol.P, vol.U, vol.phase, out, work, vol.fraction, vol.region)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            plan.run_pass(p, vol.P, vol.U, vol.phase, out, work, vol.fraction, vol.region)
        e1.record(stream)
        e1.synchronize()
        res.append(e0.elapsed_time(e1) / steps)
    return res


def cpu_baseline(w, crop):
    """The FP64 NumPy oracle (one core) on a crop of the same workload: voxels/s."""
    import numpy as np
    sys.path.insert(0, str(ROOT / "oracle"))
    from qpd_oracle import diffraction_fp64
    from qpop_diffraction.workloads import box, host_inputs, Workload
    nx, ny, nz = w.dims
    cx, cy = min(crop, nx), min(crop, ny)
    P = w.P.reshape(3, nz, ny, nx)[:, :, :cy, :cx]
    U = w.U.reshape(3, nz, ny, nx)[:, :, :cy, :cx]
    ph = w.phase.reshape(nz, ny, nx)[:, :cy, :cx]
    sub = Workload(box(cx, cy, nz, w.params.dx, w.params.nPhase), w.atoms, (cx, cy, nz), (0, 1, 2),
                   P.reshape(3, -1), U.reshape(3, -1), ph.reshape(-1), "crop")
    oPhase, Pc, Uc, region = host_inputs(sub)
    t0 = time.perf_counter()
    diffraction_fp64(sub.params, w.atoms, oPhase, Pc, Uc, region)
    dt = time.perf_counter() - t0
    n = cx * cy * nz
    return {"value": n / dt, "unit": "voxels/s", "cores": 1, "kind": "port",
            "sample": f"FP64 NumPy oracle (restatement of the reference diffraction()) on the "
                      f"{cx}x{cy}x{nz} corner crop of the same volume, {dt:.1f} s; host has {os.cpu_count()} "
                      f"cores, numpy.fft and the elementwise steps use 1"}


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU: a volume, a batch or a served job runs whole on "
                                   "one MI355X; --gpus must be 1", "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU"}))
        sys.exit(1)
    from qpop_diffraction import Engine, Volume
    from qpop_diffraction import workloads

    nx, ny, nz = args.grid
    n = nx * ny * nz
    log(f"generating two-phase {nx}x{ny}x{nz} on the GPU")
    w = workloads.two_phase(nx, ny, nz, precision=32)
    engine = Engine(w.params, w.atoms)
    vol = Volume(w.dims, w.axis, w.P, w.U, w.phase, precision=32)
    plan = engine.plan(vol.dims, vol.axis, 32)
    work = engine.workspace(plan)
    out = torch.empty(n, dtype=torch.float64, device="cuda")

    log(f"timing FP32: {args.warmup} warmup + {args.steps} steps")
    sec = time_steps(lambda: engine.intensity(vol, out), args.steps, args.warmup)
    I32 = out.clone()
    log(f"FP32 {sec * 1e3:.3f} ms/volume")
    passes = pass_times(plan, vol, out, work, max(5, args.steps))
    bpc = [plan.pass_bytes_per_cell(p, vol.P, vol.U, vol.phase) for p in range(3)]
    gbs = [bpc[p] * n / (passes[p] * 1e-3) / 1e9 for p in range(3)]
    dom = int(np.argmax(passes))
    log(f"passes ms {passes}, GB/s {gbs}")

    result = {
        "metric": "voxels/s at 512x512x104, FP32 (ms/volume FP32 and FP64 beside it)",
        "value": n / sec,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": f"two-phase Voronoi volume {nx}x{ny}x{nz}, 40 grains, 5-atom PbTiO3-like cell, "
                               f"hkl 002, one volume per step", "grid": [nx, ny, nz], "precision": "fp32",
                   "launches_per_volume": plan.launches},
        "ms_per_volume_fp32": sec * 1e3,
        "kernels": [{"pass": p, "ms": passes[p], "bytes_per_cell": bpc[p], "GB_s": gbs[p]} for p in range(3)],
        "roofline": {"bound": "hbm", "kernel": f"pass {dom}", "achieved": gbs[dom], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": gbs[dom] / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": bpc[dom] * n},
    }
    if not args.no_fp64:
        del vol
        w64 = workloads.two_phase(nx, ny, nz, precision=64)
        vol64 = Volume(w64.dims, w64.axis, w64.P, w64.U, w64.phase, precision=64)
        out64 = torch.empty(n, dtype=torch.float64, device="cuda")
        log("timing FP64")
        sec64 = time_steps(lambda: engine.intensity(vol64, out64), max(3, args.steps // 4), 1)
        peak = out64.max()
        mask = out64 >= 1e-10 * peak
        dl = (torch.log10(I32[mask]) - torch.log10(out64[mask])).abs().max().item()
        relpk = ((I32 - out64).abs().max() / peak).item()
        result["ms_per_volume_fp64"] = sec64 * 1e3
        result["accuracy_fp32_vs_fp64"] = {"max_abs_dlog10_I": dl, "max_abs_dI_over_peak": relpk,
                                           "floor": "voxels >= 1e-10 of the peak"}
        del vol64, out64, w64
    if args.cpu_crop > 0:
        log("CPU baseline")
        result["cpu_baseline"] = cpu_baseline(w, args.cpu_crop)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
