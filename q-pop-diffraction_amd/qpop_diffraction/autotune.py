"""Autotuner of the launch schedules of a plan.

Each pass of a libqpd plan can run several schedules (qpd_plan_schedules): the
compile-time kernel of its length with different tile widths, a long strided length
split into two launches N1 x N2 in several ways, or the generic run-time-radix kernel.
The reference leaves this choice to cuFFT / HeFFTe plan creation
(src/heffte/qpop_diffraction_mgpu.cu:154-157); here each candidate is timed on the
device with HIP events on the stream it runs on, the fastest kept, and the choice
remembered by name in a JSON cache keyed by device, precision, grid layout and atom
classes, so that later plans of the same shape start tuned.

Timing a pass reruns it on the workspace left by the passes before it; a split pass
transforms its inputs in place, so the workspace holds garbage afterwards and the
caller recomputes the volume (``tune`` does, when it is given the output buffer).
"""

from __future__ import annotations

import hashlib
import json
import os
from pathlib import Path

import numpy as np


def default_cache_path():
    env = os.environ.get("QPD_TUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "qpop_diffraction" / "schedules.json"


class ScheduleCache:
    """Chosen schedule names per (device, plan shape); a JSON file, or memory only when path is None."""

    def __init__(self, path="default"):
        self.path = default_cache_path() if path == "default" else (Path(path) if path else None)
        self.entries = {}
        if self.path is not None and self.path.exists():
            try:
                self.entries = json.loads(self.path.read_text())
            except (OSError, ValueError):
                self.entries = {}

    def get(self, key):
        return self.entries.get(key)

    def put(self, key, value):
        self.entries[key] = value
        if self.path is None:
            return
        try:
            self.path.parent.mkdir(parents=True, exist_ok=True)
            tmp = self.path.with_suffix(".tmp")
            tmp.write_text(json.dumps(self.entries, indent=1, sort_keys=True))
            os.replace(tmp, self.path)
        except OSError:
            pass  # a read-only home: keep the choice in memory


def device_name():
    import torch
    props = torch.cuda.get_device_properties(torch.cuda.current_device())
    return f"{props.name}/{getattr(props, 'gcnArchName', '')}"


def plan_key(plan, fields_kind=""):
    """Device, precision, layout, pass order and arithmetic, and the atom positions (which set
    the classes of each pass)."""
    h = hashlib.sha1(np.round(plan._R, 12).tobytes()).hexdigest()[:12]
    return (f"{device_name()}|fp{plan.precision}|dims{'x'.join(map(str, plan.dims))}|"
            f"axis{''.join(map(str, plan.axis))}|passes{''.join(map(str, plan.pass_dims))}|"
            f"bits{'-'.join(map(str, plan.pass_bits))}-{plan.storage_bits}|atoms{plan.n_atom}x{plan.n_phase}:{h}{fields_kind}")


def time_pass(plan, which, P, U, phase, out, work, fraction=None, region=None, reps=3, warmup=1):
    """Mean milliseconds of one run of a pass, HIP events on the current stream."""
    import torch
    st = torch.cuda.current_stream()
    for _ in range(warmup):
        plan.run_pass(which, P, U, phase, out, work, fraction, region)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        plan.run_pass(which, P, U, phase, out, work, fraction, region)
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(plan, P, U, phase, out, work, fraction=None, region=None, reps=3, cache=None, force=False):
    """Select the fastest schedule of every pass of plan (in place); returns a report dict.

    P, U, phase (and fraction, region) are the device fields of a representative
    volume, out its intensity buffer and work a workspace of plan.workspace_bytes.
    With a cache, a stored choice is applied without timing unless force is set.
    The volume is recomputed with the chosen schedules before returning.
    """
    kind = ("|frac" if fraction is not None else "") + ("|region" if region is not None else "")
    key = plan_key(plan, kind)
    names = [plan.schedules(d) for d in range(3)]
    report = {"key": key, "passes": []}
    stored = cache.get(key) if cache is not None and not force else None
    if stored is not None and all(stored.get(str(d)) in names[d] for d in range(3)):
        for d in range(3):
            plan.set_schedule(d, names[d].index(stored[str(d)]))
        report["cached"] = True
        report["passes"] = [{"pass": d, "chosen": stored[str(d)]} for d in range(3)]
        plan.intensity(P, U, phase, out, work, fraction, region)
        return report
    # valid inputs for every pass: one volume with the current schedules
    plan.intensity(P, U, phase, out, work, fraction, region)
    choice = {}
    for d in range(3):
        if len(names[d]) == 1:
            choice[str(d)] = names[d][0]
            report["passes"].append({"pass": d, "chosen": names[d][0], "ms": None, "candidates": 1})
            continue
        timings = []
        for i in range(len(names[d])):
            plan.set_schedule(d, i)
            timings.append(time_pass(plan, d, P, U, phase, out, work, fraction, region, reps=reps))
            # the passes after d read what d wrote; a split step also rewrote d's inputs in place
            plan.intensity(P, U, phase, out, work, fraction, region)
        best = int(np.argmin(timings))
        plan.set_schedule(d, best)
        plan.intensity(P, U, phase, out, work, fraction, region)
        choice[str(d)] = names[d][best]
        report["passes"].append({"pass": d, "chosen": names[d][best], "ms": timings[best],
                                 "default_ms": timings[0], "candidates": len(names[d]),
                                 "all": dict(zip(names[d], timings))})
    if cache is not None:
        cache.put(key, choice)
    report["cached"] = False
    return report
