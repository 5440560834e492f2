#!/usr/bin/env python3
"""Benchmark of the kinematic-diffraction path on one MI355X.

Workload: a 512 x 512 x 104 two-phase volume (40 periodic Voronoi grains,
tetragonal P = +-0.6 z with c/a = 1.06 through the u_z ramp, or rhombohedral
P = 0.4 (+-1, +-1, +-1)/sqrt 3; u = 0.5 % in-plane ramp + Gaussian random field of
rms 0.005 nm), PbTiO3-like 5-atom cell, reflection 002, synthetic seeded data
generated on the GPU.  One step = one volume: source, 3-D transforms and the
fftshifted intensity (three kernel launches), inputs resident in HBM.

Prints one JSON line: voxels/s of the FP32 plan (FP32 inputs and spectra, FP64
transforms: the default arithmetic) as ``value``, ms per volume, the same volume
with FP32 transforms and with FP64 throughout beside it, |dlog10 I| of each
against FP64, per-pass kernel times and GB/s, the roofline entry of the dominant
kernel, the batched, served and 4000 x 4000 x 68 legs (every batched and served
job against its FP64 result), and the CPU baseline (the FP64 NumPy oracle on the
same 512 x 512 x 104 volume, one core).

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
    ap.add_argument("--cpu-crop", type=int, default=512, help="lateral size of the CPU-baseline crop (0: skip)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 48 snapshots x 4 reflections leg")
    ap.add_argument("--batch-streams", type=int, default=4)
    ap.add_argument("--no-serve", action="store_true", help="skip the served Poisson-stream leg")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-rate", type=float, default=20.0, help="mean arrivals per second")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch schedules")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the 4000 x 4000 x 68 polar-vortex volume (about 140 GB of HBM) and its FP64 crop")
    ap.add_argument("--large-grid", type=int, nargs=3, default=(4000, 4000, 68), metavar=("NX", "NY", "NZ"))
    ap.add_argument("--large-only", action="store_true",
                    help="run only the large-volume leg (for profiling it); prints its JSON object")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first copy BYTES device to device three times (a kernel of known traffic for counters)")
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
    """The FP64 NumPy oracle (one core) on crops of the same workload: voxels/s.

    `value` is the lateral crop x crop x nz corner; `grids` adds the 64x64x32 grid of the
    CPU-runnable case and the two smaller grids of the served mix (128x128x32, 256x256x64),
    each the corner crop of the same volume."""
    import numpy as np
    sys.path.insert(0, str(ROOT / "oracle"))
    from qpd_oracle import diffraction_fp64
    from qpop_diffraction.workloads import box, host_inputs, Workload
    nx, ny, nz = w.dims

    def timed(cx, cy, cz):
        P = w.P.reshape(3, nz, ny, nx)[:, :cz, :cy, :cx]
        U = w.U.reshape(3, nz, ny, nx)[:, :cz, :cy, :cx]
        ph = w.phase.reshape(nz, ny, nx)[:cz, :cy, :cx]
        sub = Workload(box(cx, cy, cz, w.params.dx, w.params.nPhase), w.atoms, (cx, cy, cz), (0, 1, 2),
                       P.reshape(3, -1), U.reshape(3, -1), ph.reshape(-1), "crop")
        oPhase, Pc, Uc, region = host_inputs(sub)
        t0 = time.perf_counter()
        diffraction_fp64(sub.params, w.atoms, oPhase, Pc, Uc, region)
        return time.perf_counter() - t0

    cx, cy = min(crop, nx), min(crop, ny)
    dt = timed(cx, cy, nz)
    grids = []
    for g in ((64, 64, 32), (128, 128, 32), (256, 256, 64)):
        if g[0] <= nx and g[1] <= ny and g[2] <= nz:
            t = timed(*g)
            grids.append({"grid": list(g), "s": t, "voxels_per_s": g[0] * g[1] * g[2] / t})
    n = cx * cy * nz
    what = (f"the whole {nx}x{ny}x{nz} volume" if (cx, cy) == (nx, ny)
            else f"the {cx}x{cy}x{nz} corner crop of the same volume")
    return {"value": n / dt, "unit": "voxels/s", "cores": 1, "kind": "port",
            "sample": f"FP64 NumPy oracle (restatement of the reference diffraction()) on {what}, {dt:.1f} s; "
                      f"host has {os.cpu_count()} cores, numpy.fft and the elementwise steps use 1",
            "grids": grids}


def max_dlog10(I, ref):
    """max |log10 I - log10 ref| over voxels >= 1e-10 of the peak, and max |dI| / peak (device tensors)."""
    import torch
    peak = ref.max()
    mask = ref >= 1e-10 * peak
    dl = (torch.log10(I[mask]) - torch.log10(ref[mask])).abs().max().item()
    return dl, ((I - ref).abs().max() / peak).item()


def batch_leg(args):
    """48 twin-stack snapshots x 4 reflections: one graph replay per batch, against one job at a
    time; the default arithmetic (FP64 transforms on FP32 spectra) and FP32 transforms, each job
    of both against its FP64 volume."""
    import torch
    from qpop_diffraction import workloads
    from qpop_diffraction.autotune import ScheduleCache
    from qpop_diffraction.batch import Batch, PlanCache, Reflection, plain_job, resident_job
    grids = workloads.batch_grids(48, seed=3)
    log(f"batch: generating 48 snapshots {grids[:4]} ...")
    snaps = [workloads.twin_snapshot(("twin", i), L, L, d, seed=100 + i) for i, (L, d) in enumerate(grids)]
    refls = [Reflection(h, workloads.pto_atoms(h)) for h in workloads.REFLECTIONS]

    def build(arith):
        t0 = time.perf_counter()
        b = Batch(snaps, refls, n_streams=args.batch_streams, graph=True, plans=PlanCache(32, arith=arith))
        t_plan = time.perf_counter() - t0
        t0 = time.perf_counter()
        b.tune(ScheduleCache(None))
        t_tune = time.perf_counter() - t0
        b.capture()
        log(f"batch ({arith}): {len(b.jobs)} jobs, {len(b.plans)} plans in {t_plan:.2f} s, tuned in {t_tune:.1f} s, "
            f"capture {b.stats.capture_s:.2f} s")
        return b, t_plan, t_tune

    b, t_plan, t_tune = build("auto")
    vox = sum(j.plan.n for j in b.jobs)
    reps = 5
    sec = time_steps(b.run, reps, 1)
    # the same jobs one at a time: resident fields, cached plans, one stream, waiting for each
    work = torch.empty(max(j.plan.workspace_bytes for j in b.jobs), dtype=torch.uint8, device="cuda")

    def serial():
        for j in b.jobs:
            resident_job(b.plans, snaps[j.snapshot], j.reflection, refls[j.reflection], j.out, work)
            torch.cuda.synchronize()
    sec_serial = time_steps(serial, 2, 1)
    del work
    # the plain form: what one drop-in call does per job (plan, buffers, intensity to host)
    plain_job(snaps[0], refls[0])
    t0 = time.perf_counter()
    for j in b.jobs:
        plain_job(snaps[j.snapshot], refls[j.reflection])
    sec_plain = time.perf_counter() - t0
    # the same batch with FP32 transforms
    b32, _, _ = build("fp32")
    sec32 = time_steps(b32.run, reps, 1)
    # accuracy: every job of both batches against its FP64 volume
    b.run()
    b32.run()
    torch.cuda.synchronize()
    worst, worst32, n_over = 0.0, 0.0, 0
    plans64 = PlanCache(64)
    for i, (L, d) in enumerate(grids):
        s64 = workloads.twin_snapshot(("twin", i), L, L, d, seed=100 + i, precision=64)
        out64 = torch.empty(L * L * d, dtype=torch.float64, device="cuda")
        w64 = None
        for r in range(len(refls)):
            plan = plans64.get(s64, r, refls[r])
            if w64 is None or w64.numel() < plan.workspace_bytes:
                w64 = torch.empty(plan.workspace_bytes, dtype=torch.uint8, device="cuda")
            resident_job(plans64, s64, r, refls[r], out64, w64)
            dl = max_dlog10(b.result(i, r), out64)[0]
            worst = max(worst, dl)
            n_over += dl > 1e-3
            worst32 = max(worst32, max_dlog10(b32.result(i, r), out64)[0])
        del s64, w64, out64
        plans64 = PlanCache(64)
    n_jobs = len(b.jobs)
    arith = sorted({f"{'-'.join(map(str, j.plan.pass_bits))}/{j.plan.storage_bits}" for j in b.jobs})
    res = {
        "workload": "48 two-variant 90-degree twin-stack snapshots (lateral 128-512 step 64, depth 24-104 from "
                    f"{list(workloads.DEPTHS)}) x reflections 002, 103, 202, 113 = {n_jobs} volumes, FP32 inputs",
        "grids": [list(g) for g in grids],
        "volumes": n_jobs, "voxels": vox, "streams": b.n_streams, "plans": len(b.plans),
        "plan_build_s": t_plan, "autotune_s": t_tune,
        "arithmetic": f"transforms/spectra bits per job: {arith}",
        "graph_replay_ms": sec * 1e3, "volumes_per_s": n_jobs / sec, "voxels_per_s": vox / sec,
        "serial_resident_volumes_per_s": n_jobs / sec_serial,
        "plain_volumes_per_s": n_jobs / sec_plain,
        "speedup_vs_plain": sec_plain / sec, "speedup_vs_serial_resident": sec_serial / sec,
        "plain_form": "one job at a time as one drop-in diffraction() call does it on resident fields: plan "
                      "built and destroyed, workspace and output allocated, float64 intensity copied to host, "
                      "synchronous",
        "max_abs_dlog10_I_vs_fp64": worst, "jobs_over_1e-3": int(n_over),
        "dlog10_sample": f"every job ({n_jobs}) against its FP64 volume",
        "fp32_transforms": {"volumes_per_s": n_jobs / sec32, "graph_replay_ms": sec32 * 1e3,
                            "max_abs_dlog10_I_vs_fp64": worst32},
    }
    log(f"batch: {res['volumes_per_s']:.1f} vol/s graph ({n_jobs / sec32:.1f} with FP32 transforms), "
        f"{res['serial_resident_volumes_per_s']:.1f} serial, {res['plain_volumes_per_s']:.1f} plain; dlog10 {worst:.2e} "
        f"({worst32:.2e} FP32 transforms)")
    del b, b32, snaps
    torch.cuda.empty_cache()
    return res


def serve_leg(args):
    """Poisson stream of jobs over 16 resident hexagonal-twin snapshots per grid, 200 GB budget."""
    import numpy as np
    import torch
    from qpop_diffraction import workloads
    from qpop_diffraction.batch import PlanCache, Reflection, plain_job, resident_job
    from qpop_diffraction.serve import Server, latency_summary, poisson_requests
    grids = workloads.SERVE_GRIDS
    refls = [Reflection(h, workloads.pto_atoms(h)) for h in workloads.REFLECTIONS]
    store = workloads.serve_store()
    log("serve: loading 16 snapshots per grid")
    store.preload([(g, s) for g in grids for s in range(16)])
    srv = Server(store, refls, n_slots=args.serve_slots, hbm_budget=200e9)
    from qpop_diffraction.autotune import ScheduleCache, tune
    cache = ScheduleCache(None)
    for g in grids:
        snap = store.acquire((g, 0))
        for r in range(len(refls)):
            plan = srv.plans.get(snap, r, refls[r])
            o = torch.empty(plan.n, dtype=torch.float64, device="cuda")
            w = torch.empty(plan.workspace_bytes, dtype=torch.uint8, device="cuda")
            tune(plan, snap.P, snap.U, snap.phase, o, w, cache=cache)
            del o, w
        store.release((g, 0))
    warm = poisson_requests(4 * args.serve_slots * len(grids), 0.0, grids, 16, 4, seed=99)
    srv.run(warm, realtime=False)
    reqs = poisson_requests(args.serve_requests, args.serve_rate, grids, 16, 4, seed=5)
    log(f"serve: {len(reqs)} requests at {args.serve_rate}/s")
    done = srv.run(reqs, realtime=True)
    online = latency_summary(done, srv.elapsed_s)
    log(f"serve: p50 {online['p50_ms']:.2f} ms p99 {online['p99_ms']:.2f} ms")
    done_c = srv.run(reqs, realtime=False)
    cap = latency_summary(done_c, srv.elapsed_s)
    t0 = time.perf_counter()
    for r in reqs:
        plain_job(store.acquire(r.key), refls[r.reflection])
        store.release(r.key)
    sec_plain = time.perf_counter() - t0
    # accuracy: every served job (as the server delivers it to host memory) against its FP64
    # volume, in a run of its own (the comparison runs in the result callback)
    plans64 = PlanCache(64)
    snaps64, acc = {}, {"worst": 0.0, "n": 0, "over": 0}

    def check(req, host):
        key = req.key
        if key not in snaps64:
            snaps64[key] = workloads.hex_snapshot(key, *req.grid, seed=1000 + int(req.snapshot), precision=64)
        s64 = snaps64[key]
        plan = plans64.get(s64, req.reflection, refls[req.reflection])
        o64 = torch.empty(plan.n, dtype=torch.float64, device="cuda")
        w = torch.empty(plan.workspace_bytes, dtype=torch.uint8, device="cuda")
        resident_job(plans64, s64, req.reflection, refls[req.reflection], o64, w)
        dl = max_dlog10(host.to("cuda").reshape(-1), o64)[0]
        acc["worst"] = max(acc["worst"], dl)
        acc["n"] += 1
        acc["over"] += dl > 1e-3

    srv_check = Server(store, refls, n_slots=args.serve_slots, hbm_budget=200e9, on_result=check, plans=srv.plans)
    srv_check.run(reqs, realtime=False)
    worst = acc["worst"]
    del snaps64
    res = {
        "workload": f"Poisson stream, mean {args.serve_rate} req/s, {len(reqs)} jobs; grid uniform from "
                    "128x128x32, 256x256x64, 512x512x104; snapshot id uniform from 16 three-variant hexagonal-twin "
                    "fields per grid (resident); reflection uniform from 002, 103, 202, 113; HBM budget 200 GB; "
                    f"{args.serve_slots} slots; latency = arrival to float64 intensity in pinned host memory",
        "online": online,
        "capacity_volumes_per_s": cap["volumes_per_s"],
        "plain_volumes_per_s": len(reqs) / sec_plain,
        "speedup_vs_plain": sec_plain / cap["elapsed_s"],
        "peak_hbm_bytes": srv.peak_bytes, "deferred_admissions": srv.deferred,
        "max_abs_dlog10_I_vs_fp64": worst, "jobs_over_1e-3": int(acc["over"]),
        "dlog10_sample": f"every served job ({acc['n']}) as delivered to host memory, against its FP64 volume",
    }
    log(f"serve: capacity {cap['volumes_per_s']:.1f} vol/s vs plain {res['plain_volumes_per_s']:.1f}")
    return res


def measured_traffic(grid, pass_index, schedule):
    """HBM bytes per launch of a kernel, from the committed rocprofv3 counter summaries
    (profiles/*_traffic.json), or None.

    A summary holds FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what the same
    counter reads for a device copy of known size; it is used only when it was taken for the
    same grid, pass and schedule.
    """
    for f in sorted((ROOT / "profiles").glob("*_traffic.json")):
        try:
            t = json.loads(f.read_text())
        except (OSError, ValueError):
            continue
        if list(t.get("grid", [])) != list(grid):
            continue
        for k in t.get("kernels", []):
            if k.get("pass") == pass_index and k.get("schedule") == schedule:
                return k.get("bytes")
    return None


def copy_ceiling(nbytes=1 << 30, reps=5):
    """Device-to-device copy of nbytes (torch copy_, one read and one write of each byte),
    timed with HIP events: the bandwidth a plain streaming kernel reaches on this GPU."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        dst.copy_(src)
    e1.record(st)
    e1.synchronize()
    sec = e0.elapsed_time(e1) / reps * 1e-3
    del src, dst
    return 2 * nbytes / sec / 1e9


def roofline_of(name, gb_s, bytes_per_launch, traffic, ceiling):
    return {"bound": "hbm", "kernel": name, "achieved": gb_s, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": gb_s / HBM_PEAK_GBS, "traffic": traffic, "bytes_per_launch": bytes_per_launch,
            "copy_ceiling": {"GB_s": ceiling, "frac": gb_s / ceiling,
                             "kernel": "device-to-device copy of 1 GiB (torch copy_), HIP events"}}


def ingest_leg(nx, ny, nz, reps=10):
    """Field ingest on the GPU (raw float32 P, float64 u in m, int32 materials over a deeper stack
    with a 16-cell substrate below the crop) against the same preparation in NumPy on the host."""
    import numpy as np
    import torch
    from qpop_diffraction.constants import L0
    from qpop_diffraction.ingest import ingest
    nz_full, z0 = nz + 16, 16
    gen = torch.Generator(device="cuda").manual_seed(11)
    P = [0.5 * torch.randn((nz, ny, nx), generator=gen, device="cuda") for _ in range(3)]
    U = [1e-11 * torch.randn((nz, ny, nx), generator=gen, device="cuda", dtype=torch.float64) for _ in range(3)]
    zz = torch.arange(nz_full, device="cuda").view(-1, 1, 1)
    mat = torch.where(zz < z0, 0, 1 + (torch.rand((nz_full, ny, nx), generator=gen, device="cuda") < 0.5).int())
    mat = mat.to(torch.int32).contiguous()
    dc = np.array([0.0, 0.012, -0.004])
    phase_of = np.array([-1, 0, 1], dtype=np.int8)
    dz = 0.39
    ingest(P, U, mat, dc, phase_of, z0, dz, 32)
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        Pd, Ud, phd = ingest(P, U, mat, dc, phase_of, z0, dz, 32)
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    n = nx * ny * nz
    nbytes = n * (3 * 4 + 3 * 8 + 3 * 4 + 3 * 8 + 1) + nx * ny * nz_full * 4
    # the same preparation in NumPy (cli.py:189-217), host arrays already in memory
    Ph = [t.cpu().numpy() for t in P]
    Uh = [t.cpu().numpy() for t in U]
    mh = mat.cpu().numpy()
    t0 = time.perf_counter()
    Pn = np.stack(Ph).astype(np.float32)
    Un = np.stack(Uh) / L0
    d = dc[mh]
    Un[2] += ((np.cumsum(d, axis=0) - 0.5 * d) * dz)[z0:]
    phn = phase_of[mh[z0:]]
    host_s = time.perf_counter() - t0
    same = (np.array_equal(Pd.cpu().numpy().reshape(3, -1), Pn.reshape(3, -1)) and
            np.array_equal(Ud.cpu().numpy().reshape(3, -1), Un.reshape(3, -1)) and
            np.array_equal(phd.cpu().numpy(), phn.reshape(-1)))
    return {"workload": f"{nx}x{ny}x{nz} crop of a {nz_full}-deep stack: float32 P, float64 u (m), int32 materials",
            "ms": ms, "GB_s": nbytes / (ms * 1e-3) / 1e9, "bytes_per_launch": nbytes,
            "frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "host_numpy_ms": host_s * 1e3, "bit_identical_to_host": bool(same)}


def output_leg(I, dims):
    """Writing the float64 intensity of the volume to diffraction.h5: copy-then-write against the
    streamed writer (slabs copied to pinned memory on their own stream while the previous one
    is written)."""
    import tempfile
    import numpy as np
    import torch
    from qpop_diffraction.io_utils import write_diffraction_output, write_diffraction_streamed
    nx, ny, nz = dims
    q = [np.linspace(-1.0, 1.0, m) for m in (nx, ny, nz)]
    with tempfile.TemporaryDirectory(prefix="qpd_out_") as d:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        host = I.cpu().numpy().reshape(nz, ny, nx)
        write_diffraction_output(os.path.join(d, "plain"), host, *q, 0, zyx=True)
        plain = time.perf_counter() - t0
        del host
        write_diffraction_streamed(os.path.join(d, "warm"), I, *q, 0)  # staging buffers, file system
        streamed = write_diffraction_streamed(os.path.join(d, "streamed"), I, *q, 0)
    nbytes = I.numel() * 8
    return {"bytes": nbytes, "copy_then_write_s": plain, "streamed_s": streamed,
            "streamed_GB_s": nbytes / streamed / 1e9, "speedup": plain / streamed,
            "target": "HDF5 file in a temporary directory of the host"}


def rsm_leg(I, dims, spacing):
    """Reciprocal-space maps (q_x-q_y, q_x-q_z, q_y-q_z) of the volume's intensity from device axis
    sums, against the reference's way (numpy.histogram2d of every voxel, on the host copy) for one
    of the three maps."""
    import numpy as np
    import torch
    sys.path.insert(0, str(ROOT / "oracle"))
    from qpd_oracle import rsm_map
    from qpop_diffraction.rsm import reciprocal_space_maps
    nx, ny, nz = dims
    q = [np.fft.fftshift(np.fft.fftfreq(m, spacing)) * 2 * np.pi for m in (nx, ny, nz)]
    reciprocal_space_maps(I, dims, (0, 1, 2), q)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    maps = reciprocal_space_maps(I, dims, (0, 1, 2), q)
    dev_s = time.perf_counter() - t0
    Ixyz = I.cpu().numpy().reshape(nz, ny, nx).transpose(2, 1, 0)
    t0 = time.perf_counter()
    ref = rsm_map(Ixyz, q, 0, 2)[2]
    host_s = time.perf_counter() - t0
    err = float(np.max(np.abs(maps["qx_qz"][2] - ref)) / np.max(ref))
    return {"maps": 3, "device_s": dev_s, "host_histogram_s_one_map": host_s,
            "max_rel_diff_qx_qz": err, "note": "device: three axis sums over the volume in HBM plus three "
            "2-D histograms of the sums; host: numpy.histogram2d of every voxel for the q_x-q_z map"}


def large_leg(args):
    """One large volume on one GPU in the default arithmetic and with FP32 transforms: time per
    volume, per-pass rates, and each one's error against FP64 on a 512 x 512 x nz crop around G."""
    import numpy as np
    import torch
    from qpop_diffraction import Engine, Volume, workloads
    from qpop_diffraction.autotune import ScheduleCache
    from qpop_diffraction.large import crop_of, fp64_crop
    nx, ny, nz = args.large_grid
    # the peak below counts this volume's fields, workspace and output only
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats()
    base = torch.cuda.memory_allocated()
    t0 = time.perf_counter()
    v = workloads.vortex(nx, ny, nz, precision=32)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    log(f"large: {v.name} generated in {t_gen:.1f} s")
    vol = Volume(v.dims, v.axis, v.P, v.U, v.phase, precision=32)
    ceiling = copy_ceiling()

    def one(arith):
        eng = Engine(v.params, v.atoms, arith=arith)
        out = torch.empty(vol.n, dtype=torch.float64, device="cuda")
        tuned = eng.tune(vol, cache=ScheduleCache(None), out=out)
        plan = eng.plan(vol.dims, vol.axis, 32)
        sec = time_steps(lambda: eng.intensity(vol, out), 3, 1)
        passes = pass_times(plan, vol, out, eng.workspace(plan), 2)
        eng.intensity(vol, out)
        crop = crop_of(out, v.dims)
        hbm = torch.cuda.max_memory_allocated() - base
        names = [p["chosen"] for p in tuned["passes"]]
        bpc = [plan.pass_bytes_per_cell(p, vol.P, vol.U, vol.phase) for p in range(3)]
        gbs = [bpc[p] * vol.n / (passes[p] * 1e-3) / 1e9 for p in range(3)]
        dom = int(np.argmax(passes))
        log(f"large {arith}: {sec * 1e3:.1f} ms/volume, passes {passes}, peak HBM {hbm / 1e9:.1f} GB")
        res = {"arithmetic": f"{'/'.join(map(str, plan.pass_bits))}-bit transforms (passes 0/1/2), "
                             f"{plan.storage_bits}-bit spectra, estimate {plan.dlog10_estimate:.1e}",
               "ms_per_volume": sec * 1e3, "voxels_per_s": vol.n / sec, "passes_ms": passes, "schedules": names,
               "GB_s": gbs, "bytes_per_cell": bpc,
               "traffic": [measured_traffic((nx, ny, nz), p, names[p]) for p in range(3)],
               "roofline": roofline_of(f"pass {dom}: {names[dom]}", gbs[dom], bpc[dom] * vol.n,
                                       measured_traffic((nx, ny, nz), dom, names[dom]), ceiling),
               "peak_hbm_bytes": hbm}
        del out
        eng._work.clear()
        torch.cuda.empty_cache()
        return res, crop, eng.tables

    res, I32, tables = one("auto")
    res_f, I32f, _ = one("fp32")
    t0 = time.perf_counter()
    ref = fp64_crop(v, v.params, tables, log=log)
    t_ref = time.perf_counter() - t0
    dl, rel = max_dlog10(I32, ref)
    dlf, relf = max_dlog10(I32f, ref)
    floor = "voxels >= 1e-10 of the crop's peak"
    res_f["accuracy_vs_fp64_crop"] = {"max_abs_dlog10_I": dlf, "max_abs_dI_over_peak": relf, "floor": floor}
    res_f["note"] = "explicit FP32 arithmetic (arith='fp32'): no accuracy bound"
    res.update({"workload": f"{v.name}, layout z fastest, FP32, reflection 002", "grid": [nx, ny, nz],
                "generate_s": t_gen,
                "accuracy_vs_fp64_crop": {"crop": [512, 512, nz], "max_abs_dlog10_I": dl, "max_abs_dI_over_peak": rel,
                                          "floor": floor, "fp64_crop_s": t_ref},
                "fp32_transforms": res_f})
    return res


def main():
    # the library and the atom-file reader print progress on stdout, as the reference
    # does; the bench keeps stdout for its one JSON line
    real_stdout = sys.stdout
    sys.stdout = sys.stderr
    try:
        result = run()
    finally:
        sys.stdout = real_stdout
    if result is not None:
        print(json.dumps(result), flush=True)


def run():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU: a volume, a batch or a served job runs whole on "
                                   "one MI355X; --gpus must be 1", "n_gpus": args.gpus}), file=sys.__stdout__)
        sys.exit(2)
    import numpy as np
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU"}), file=sys.__stdout__)
        sys.exit(1)
    if args.copy_calibration > 0:
        src = torch.ones(args.copy_calibration // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: 3 x {args.copy_calibration} bytes")
        del src, dst
    if args.large_only:
        return {"large": large_leg(args)}
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
    from qpop_diffraction.autotune import ScheduleCache
    if args.no_tune:
        tuned = {"passes": [{"pass": d, "chosen": plan.schedules(d)[plan.schedule(d)]} for d in range(3)]}
    else:
        tuned = engine.tune(vol, cache=ScheduleCache(None), out=out)
    log("schedules: " + "; ".join(f"pass {p['pass']}: {p['chosen']}" for p in tuned["passes"]))

    log(f"timing FP32 inputs (default arithmetic: {plan.compute_bits}-bit transforms, {plan.storage_bits}-bit "
        f"spectra): {args.warmup} warmup + {args.steps} steps")
    sec = time_steps(lambda: engine.intensity(vol, out), args.steps, args.warmup)
    I32 = out.clone()
    log(f"FP32 inputs {sec * 1e3:.3f} ms/volume")
    # the same volume with FP32 transforms (explicit arithmetic, no accuracy bound)
    eng_f = Engine(w.params, w.atoms, arith="fp32")
    out_f = torch.empty(n, dtype=torch.float64, device="cuda")
    if not args.no_tune:
        eng_f.tune(vol, cache=ScheduleCache(None), out=out_f)
    sec_f = time_steps(lambda: eng_f.intensity(vol, out_f), args.steps, args.warmup)
    plan_f = eng_f.plan(vol.dims, vol.axis, 32)
    passes_f = pass_times(plan_f, vol, out_f, eng_f.workspace(plan_f), max(5, args.steps))
    bpc_f = [plan_f.pass_bytes_per_cell(p, vol.P, vol.U, vol.phase) for p in range(3)]
    gbs_f = [bpc_f[p] * n / (passes_f[p] * 1e-3) / 1e9 for p in range(3)]
    dom_f = int(np.argmax(passes_f))
    names_f = [plan_f.schedules(p)[plan_f.schedule(p)] for p in range(3)]
    eng_f.intensity(vol, out_f)
    I32f = out_f.clone()
    del out_f, eng_f
    log(f"FP32 transforms {sec_f * 1e3:.3f} ms/volume, passes {passes_f}")
    passes = pass_times(plan, vol, out, work, max(5, args.steps))
    bpc = [plan.pass_bytes_per_cell(p, vol.P, vol.U, vol.phase) for p in range(3)]
    gbs = [bpc[p] * n / (passes[p] * 1e-3) / 1e9 for p in range(3)]
    dom = int(np.argmax(passes))
    log(f"passes ms {passes}, GB/s {gbs}")
    ceiling = copy_ceiling()
    log(f"copy ceiling {ceiling:.0f} GB/s")

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
        # the arithmetic of the default plan: FP64 transforms where the plan runs them, FP32 spectra
        "dtype": ("f64 transforms, f32 spectra" if all(b == 64 for b in plan.pass_bits)
                  else "f32/f64 transforms (" + "/".join(map(str, plan.pass_bits)) + "-bit), f32 spectra"),
        "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": f"two-phase Voronoi volume {nx}x{ny}x{nz}, 40 grains, 5-atom PbTiO3-like cell, "
                               f"hkl 002, one volume per step", "grid": [nx, ny, nz], "precision": "fp32",
                   "arithmetic": f"FP32 inputs and spectra between passes, "
                                 f"{'/'.join(map(str, plan.pass_bits))}-bit transforms (passes 0/1/2) "
                                 f"(the default of an FP32 plan)",
                   "pass_dims": list(plan.pass_dims), "launches_per_volume": plan.launches},
        "ms_per_volume_fp32": sec * 1e3,
        "fp32_transforms": {"ms_per_volume": sec_f * 1e3, "voxels_per_s": n / sec_f, "passes_ms": passes_f,
                            "pass_dims": list(plan_f.pass_dims), "schedules": names_f,
                            "roofline": roofline_of(f"pass {dom_f}: {names_f[dom_f]}", gbs_f[dom_f],
                                                    bpc_f[dom_f] * n,
                                                    measured_traffic((nx, ny, nz), dom_f, names_f[dom_f]), ceiling),
                            "note": "explicit FP32 arithmetic (arith='fp32'): no accuracy bound"},
        "kernels": [{"pass": p, "ms": passes[p], "bytes_per_cell": bpc[p], "GB_s": gbs[p],
                     "schedule": tuned["passes"][p]["chosen"]} for p in range(3)],
        "roofline": roofline_of(f"pass {dom}: {tuned['passes'][dom]['chosen']}", gbs[dom], bpc[dom] * n,
                                measured_traffic((nx, ny, nz), dom, tuned["passes"][dom]["chosen"]), ceiling),
        "traffic": [measured_traffic((nx, ny, nz), p, tuned["passes"][p]["chosen"]) for p in range(3)],
    }
    log("output path")
    result["output"] = output_leg(out, (nx, ny, nz))
    log(f"output: copy-then-write {result['output']['copy_then_write_s']:.3f} s, streamed "
        f"{result['output']['streamed_s']:.3f} s")
    log("reciprocal-space maps")
    result["rsm"] = rsm_leg(out, (nx, ny, nz), w.params.dx)
    log(f"rsm: device {result['rsm']['device_s'] * 1e3:.1f} ms for 3 maps, host histogram "
        f"{result['rsm']['host_histogram_s_one_map']:.2f} s for 1")
    del vol, out
    engine._work.clear()
    torch.cuda.empty_cache()
    if not args.no_fp64:
        w64 = workloads.two_phase(nx, ny, nz, precision=64)
        vol64 = Volume(w64.dims, w64.axis, w64.P, w64.U, w64.phase, precision=64)
        out64 = torch.empty(n, dtype=torch.float64, device="cuda")
        if not args.no_tune:
            engine.tune(vol64, cache=ScheduleCache(None), out=out64)
        log("timing FP64")
        sec64 = time_steps(lambda: engine.intensity(vol64, out64), max(3, args.steps // 4), 1)
        dl, relpk = max_dlog10(I32, out64)
        dlf, relpkf = max_dlog10(I32f, out64)
        result["ms_per_volume_fp64"] = sec64 * 1e3
        result["accuracy_fp32_vs_fp64"] = {"max_abs_dlog10_I": dl, "max_abs_dI_over_peak": relpk,
                                           "floor": "voxels >= 1e-10 of the peak",
                                           "estimate": plan.dlog10_estimate}
        result["fp32_transforms"]["accuracy_vs_fp64"] = {"max_abs_dlog10_I": dlf, "max_abs_dI_over_peak": relpkf}
        del vol64, out64, w64
        engine._work.clear()
        torch.cuda.empty_cache()
    log("field ingest")
    result["ingest"] = ingest_leg(nx, ny, nz)
    log(f"ingest: {result['ingest']['ms']:.3f} ms ({result['ingest']['GB_s']:.0f} GB/s) vs NumPy "
        f"{result['ingest']['host_numpy_ms']:.0f} ms, identical {result['ingest']['bit_identical_to_host']}")
    if not args.no_batch:
        result["batched"] = batch_leg(args)
    if not args.no_serve:
        result["served"] = serve_leg(args)
    if not args.no_large:
        torch.cuda.empty_cache()
        result["large"] = large_leg(args)
    if args.cpu_crop > 0:
        log("CPU baseline")
        result["cpu_baseline"] = cpu_baseline(w, args.cpu_crop)
    return result


if __name__ == "__main__":
    main()
