"""Batched volumes: many snapshots x reflections in one captured launch sequence.

The reference computes one volume per ``diffraction()`` call (src/python/diffraction.py:62-113,
called once by cli.py:221); a study over snapshots and reflections calls it in a loop, and
every call rebuilds its tables, moves the fields and the intensity over PCIe and waits.

Here the fields of every snapshot stay in HBM, one plan per (grid, reflection) is built once
and kept, and all jobs are enqueued on a few default-priority streams (longest job first onto
the least-loaded stream) with one workspace per stream.  The whole sequence is captured once
into a HIP graph (torch.cuda.CUDAGraph, the streams forked from and joined back into the
capture stream) and each batch is one graph replay.  Results stay in HBM, one float64 tensor
per job, laid out like the snapshot (x fastest, fftshifted).

``plain_job`` is the one-job-at-a-time form the batch is measured against: what the drop-in
``diffraction()`` does per call on fields that are already resident (plan and workspace made
for the call, intensity copied to host memory, synchronous).
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from ._native import Plan, current_stream
from .diffraction import crystal_tables


@dataclass
class Reflection:
    """A crystal with its pattern centred on one reciprocal lattice point."""
    hkl: tuple
    atoms: object
    tables: dict = field(default_factory=dict)

    def crystal(self, params):
        """(R, f, su, G, qC, V) for a run; qC depends on the run's average strain."""
        key = tuple(float(v) for v in params.strainAvg)
        if key not in self.tables:
            self.tables[key] = crystal_tables(params, self.atoms)
        return self.tables[key]


class PlanCache:
    """libqpd plans keyed by (memory dims, axis, spacing, reflection, precision).

    order and arith are the plan options of ``Plan`` (pass order, arithmetic of FP32 plans)."""

    def __init__(self, precision=32, order="auto", arith="auto"):
        self.precision = precision
        self.order, self.arith = order, arith
        self._plans = {}
        self.built = 0

    def get(self, snap, refl_index, refl):
        p = snap.params
        h = (p.dx, p.dy, p.dz)
        key = (tuple(snap.dims), tuple(snap.axis), h, tuple(float(v) for v in p.strainAvg), refl_index,
               self.precision)
        plan = self._plans.get(key)
        if plan is None:
            R, f, su, G, qC, V = refl.crystal(p)
            plan = Plan(snap.dims, snap.axis, h, R, f, su, G, qC, V, self.precision, order=self.order,
                        arith=self.arith if self.precision == 32 else "auto")
            self._plans[key] = plan
            self.built += 1
        return plan

    def __len__(self):
        return len(self._plans)


@dataclass
class Job:
    snapshot: int
    reflection: int
    plan: object = None
    out: object = None
    stream: int = 0
    cost: float = 0.0


@dataclass
class BatchStats:
    jobs: int
    streams: int
    plans: int
    workspace_bytes: int
    output_bytes: int
    capture_s: float = 0.0
    per_stream_cost: list = field(default_factory=list)


def assign_streams(costs, n_streams):
    """Longest job first onto the least-loaded stream; returns (stream of each job, loads)."""
    order = sorted(range(len(costs)), key=lambda j: -costs[j])
    load = [0.0] * n_streams
    where = [0] * len(costs)
    for j in order:
        s = int(np.argmin(load))
        where[j] = s
        load[s] += costs[j]
    return where, load


class Batch:
    """All (snapshot, reflection) jobs of a study, run as one graph replay.

    snapshots: objects with params, dims, axis, P, U, phase (device tensors) and
    optional fraction/region; reflections: list of Reflection.  jobs: list of
    (snapshot index, reflection index), default every pair.
    """

    def __init__(self, snapshots, reflections, jobs=None, precision=32, n_streams=4, graph=True, plans=None,
                 tune=False, tune_cache=None):
        import torch
        self.snapshots = snapshots
        self.reflections = reflections
        self.precision = precision
        self.plans = plans if plans is not None else PlanCache(precision)
        pairs = jobs if jobs is not None else [(s, r) for s in range(len(snapshots)) for r in range(len(reflections))]
        self.jobs = []
        for s, r in pairs:
            snap = snapshots[s]
            plan = self.plans.get(snap, r, reflections[r])
            self.jobs.append(Job(s, r, plan, cost=float(plan.n)))
        self.n_streams = max(1, min(int(n_streams), len(self.jobs)))
        where, load = assign_streams([j.cost for j in self.jobs], self.n_streams)
        for j, s in zip(self.jobs, where):
            j.stream = s
        self.streams = [torch.cuda.Stream() for _ in range(self.n_streams)]
        ws = [1] * self.n_streams
        for j in self.jobs:
            ws[j.stream] = max(ws[j.stream], j.plan.workspace_bytes)
        self.work = [torch.empty(b, dtype=torch.uint8, device="cuda") for b in ws]
        for j in self.jobs:
            j.out = torch.empty(j.plan.n, dtype=torch.float64, device="cuda")
        self.stats = BatchStats(len(self.jobs), self.n_streams, len(self.plans), sum(ws),
                                sum(j.out.numel() * 8 for j in self.jobs), per_stream_cost=load)
        self.graph = None
        self.use_graph = bool(graph)
        self.tuning = []
        if tune:
            self.tune(tune_cache)

    def tune(self, cache=None):
        """Autotune the schedules of every distinct plan on its first job (before capture)."""
        from .autotune import tune
        seen = set()
        for j in self.jobs:
            if id(j.plan) in seen:
                continue
            seen.add(id(j.plan))
            snap = self.snapshots[j.snapshot]
            self.tuning.append(tune(j.plan, snap.P, snap.U, snap.phase, j.out, self.work[j.stream],
                                    getattr(snap, "fraction", None), getattr(snap, "region", None), cache=cache))
        return self.tuning

    def _enqueue(self):
        import torch
        base = torch.cuda.current_stream()
        for s in self.streams:
            s.wait_stream(base)
        for j in self.jobs:
            snap = self.snapshots[j.snapshot]
            st = self.streams[j.stream]
            j.plan.intensity(snap.P, snap.U, snap.phase, j.out, self.work[j.stream],
                             getattr(snap, "fraction", None), getattr(snap, "region", None),
                             stream=_handle(st))
        for s in self.streams:
            base.wait_stream(s)

    def capture(self):
        """Run the sequence once eagerly, then capture it into a HIP graph."""
        import torch
        self._enqueue()
        torch.cuda.synchronize()
        if not self.use_graph:
            return
        t0 = time.perf_counter()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, capture_error_mode="thread_local"):
            self._enqueue()
        torch.cuda.synchronize()
        self.graph = g
        self.stats.capture_s = time.perf_counter() - t0

    def run(self):
        """Enqueue one batch on the current stream (graph replay, or eager when not captured)."""
        if self.graph is not None:
            self.graph.replay()
        else:
            self._enqueue()

    def result(self, snapshot, reflection):
        for j in self.jobs:
            if j.snapshot == snapshot and j.reflection == reflection:
                return j.out
        raise KeyError((snapshot, reflection))


def _handle(stream):
    import ctypes
    return ctypes.c_void_p(stream.cuda_stream)


def resident_job(plans, snap, refl_index, refl, out, work):
    """One job on the current stream with a cached plan and a caller-owned workspace."""
    plan = plans.get(snap, refl_index, refl)
    plan.intensity(snap.P, snap.U, snap.phase, out, work, getattr(snap, "fraction", None),
                   getattr(snap, "region", None), stream=current_stream())
    return out


def plain_job(snap, refl, precision=32):
    """The one-job-at-a-time form: per-call plan, workspace and output, result in host memory.

    Mirrors one drop-in ``diffraction()`` call (diffraction.py:62-113) on fields that are
    already in HBM; returns the host float64 intensity.
    """
    import torch
    p = snap.params
    R, f, su, G, qC, V = refl.crystal(p)
    plan = Plan(snap.dims, snap.axis, (p.dx, p.dy, p.dz), R, f, su, G, qC, V, precision)
    try:
        work = torch.empty(max(1, plan.workspace_bytes), dtype=torch.uint8, device="cuda")
        out = torch.empty(plan.n, dtype=torch.float64, device="cuda")
        plan.intensity(snap.P, snap.U, snap.phase, out, work, getattr(snap, "fraction", None),
                       getattr(snap, "region", None))
        host = out.cpu()
    finally:
        plan.close()
    return host
