"""Serving diffraction jobs: continuous batching under an HBM budget.

A job is (grid, snapshot id, reflection).  The reference has no server: each volume is one
synchronous ``diffraction()`` call (src/python/diffraction.py:62-113) that moves the fields to
the device and the intensity back.  Here:

  * snapshot fields live in a ``SnapshotStore`` (loaded on first use, kept in HBM across
    reflection requests for the same snapshot, least-recently-used eviction when the budget
    needs room, never while a job in flight reads them);
  * a job runs in a slot: one default-priority stream with its own workspace, device output
    and pinned host buffer, cached per grid;
  * the scheduler admits queued jobs in arrival order as soon as a slot is free and the HBM
    footprint (resident snapshots + slot buffers) stays within the budget, so jobs that arrive
    while others are on the device start at once and overlap with them;
  * a job ends when its intensity is in host memory (device-to-host copy on the slot stream);
    latency is arrival to that point.

Results depend only on the job: the same plan and kernels run whatever else is in flight, so a
job's intensity is bit-identical alone and in a mix.
"""

from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np

from .batch import PlanCache, _handle


@dataclass
class Request:
    rid: int
    t_arrival: float  # seconds after the start of the stream
    grid: tuple       # (nx, ny, nz)
    snapshot: int
    reflection: int

    @property
    def key(self):
        return (tuple(self.grid), int(self.snapshot))


@dataclass
class Completion:
    rid: int
    latency_s: float
    queue_s: float
    n: int
    slot: int


def poisson_requests(n, rate, grids, n_snapshots, n_reflections, seed=0):
    """n jobs with exponential inter-arrival times of mean 1/rate; grid, snapshot, reflection uniform."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n)) if rate > 0 else np.zeros(n)
    g = rng.integers(0, len(grids), size=n)
    s = rng.integers(0, n_snapshots, size=n)
    r = rng.integers(0, n_reflections, size=n)
    return [Request(i, float(t[i]), tuple(grids[g[i]]), int(s[i]), int(r[i])) for i in range(n)]


def tensor_bytes(snap):
    ts = [snap.P, snap.U, snap.phase, getattr(snap, "fraction", None), getattr(snap, "region", None)]
    return sum(t.numel() * t.element_size() for t in ts if t is not None)


class SnapshotStore:
    """Device-resident snapshot fields keyed by (grid, id).

    spec(key) describes a snapshot without loading it (an object with params, dims, axis
    and nbytes: enough to build its plan and to budget it); loader(key) makes its fields
    on the current stream.
    """

    def __init__(self, spec, loader):
        self.spec = spec
        self.loader = loader
        self._snaps = collections.OrderedDict()  # key -> snapshot, least recently used first
        self._bytes = {}
        self._users = collections.Counter()
        self.loads = 0
        self.evictions = 0

    @property
    def nbytes(self):
        return sum(self._bytes.values())

    def resident(self, key):
        return key in self._snaps

    def footprint(self, key):
        """Bytes that acquiring key would add."""
        return 0 if key in self._snaps else int(self.spec(key).nbytes)

    def acquire(self, key):
        if key not in self._snaps:
            snap = self.loader(key)
            self._snaps[key] = snap
            self._bytes[key] = tensor_bytes(snap)
            self.loads += 1
        self._snaps.move_to_end(key)
        self._users[key] += 1
        return self._snaps[key]

    def release(self, key):
        self._users[key] -= 1

    def evict_one(self, protect=None):
        """Drop the least recently used snapshot that no job in flight reads; returns bytes freed."""
        for key in self._snaps:
            if key != protect and self._users[key] <= 0:
                del self._snaps[key]
                self.evictions += 1
                return self._bytes.pop(key)
        return 0

    def preload(self, keys):
        for k in keys:
            self.acquire(k)
            self.release(k)


class Slot:
    """One stream with cached per-grid buffers: workspace, device output, pinned host output."""

    def __init__(self, index):
        import torch
        self.index = index
        self.stream = torch.cuda.Stream()
        self.buffers = {}  # grid -> (work, out, host)
        self.job = None
        self.event = None

    def buffer_bytes(self, grid=None):
        items = self.buffers.items() if grid is None else [(grid, self.buffers[grid])] if grid in self.buffers else []
        return sum(w.numel() + o.numel() * 8 for _, (w, o, h) in items)

    def extra_bytes(self, grid, work_bytes, n):
        """Device bytes that ensure(grid, work_bytes, n) would add."""
        if grid not in self.buffers:
            return max(1, work_bytes) + 8 * n
        return max(0, work_bytes - self.buffers[grid][0].numel())

    def ensure(self, grid, work_bytes, n):
        """The grid's buffers, the workspace grown if this plan needs more than the cached one."""
        import torch
        if grid not in self.buffers:
            self.buffers[grid] = (torch.empty(max(1, work_bytes), dtype=torch.uint8, device="cuda"),
                                  torch.empty(n, dtype=torch.float64, device="cuda"),
                                  torch.empty(n, dtype=torch.float64, pin_memory=True))
        elif self.buffers[grid][0].numel() < work_bytes:
            _, out, host = self.buffers[grid]
            self.buffers[grid] = (torch.empty(work_bytes, dtype=torch.uint8, device="cuda"), out, host)
        return self.buffers[grid]

    def trim(self, keep=None):
        freed = 0
        for g in list(self.buffers):
            if g != keep:
                freed += self.buffer_bytes(g)
                del self.buffers[g]
        return freed


class Server:
    """Continuous-batching scheduler over ``n_slots`` streams with an HBM budget in bytes.

    store: a SnapshotStore; reflections: list of batch.Reflection.  on_result(request,
    host_tensor) is called when a job's intensity is in host memory; the tensor is the
    slot's pinned buffer and is reused after the call returns.
    """

    def __init__(self, store, reflections, n_slots=4, hbm_budget=200e9, precision=32, on_result=None,
                 plans=None):
        self.store = store
        self.reflections = reflections
        self.slots = [Slot(i) for i in range(n_slots)]
        self.budget = float(hbm_budget)
        self.plans = plans if plans is not None else PlanCache(precision)
        self.on_result = on_result
        self.peak_bytes = 0
        self.deferred = 0  # admission attempts held back by the budget
        self.elapsed_s = 0.0

    def used_bytes(self):
        return self.store.nbytes + sum(s.buffer_bytes() for s in self.slots)

    def plan_of(self, req):
        return self.plans.get(self.store.spec(req.key), req.reflection, self.reflections[req.reflection])

    def _need(self, req, slot, plan):
        return self.store.footprint(req.key) + slot.extra_bytes(req.grid, plan.workspace_bytes, plan.n)

    def _fits(self, req, slot, plan):
        """Whether the job fits the budget in this slot, freeing idle buffers and unused snapshots if needed."""
        if self.used_bytes() + self._need(req, slot, plan) <= self.budget:
            return True
        for s in self.slots:
            if s.job is None:
                s.trim(keep=req.grid if s is slot else None)
        while self.used_bytes() + self._need(req, slot, plan) > self.budget:
            if self.store.evict_one(protect=req.key) == 0:
                return False
        return True

    def _start(self, req, slot, plan, t0):
        import torch
        snap = self.store.acquire(req.key)
        work, out, host = slot.ensure(req.grid, plan.workspace_bytes, plan.n)
        slot.stream.wait_stream(torch.cuda.current_stream())  # a snapshot just loaded there
        plan.intensity(snap.P, snap.U, snap.phase, out, work, getattr(snap, "fraction", None),
                       getattr(snap, "region", None), stream=_handle(slot.stream))
        with torch.cuda.stream(slot.stream):
            host.copy_(out, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(slot.stream)
        slot.job = (req, time.perf_counter() - t0, host)
        slot.event = ev
        self.peak_bytes = max(self.peak_bytes, self.used_bytes())

    def run(self, requests, realtime=True, poll_s=50e-6):
        """Serve the requests; returns the completions in finishing order.

        realtime: a request is queued at its arrival time; otherwise every request is
        queued at the start (closed loop, for capacity; latency then counts from the start).
        """
        reqs = sorted(requests, key=lambda r: r.t_arrival)
        queue = collections.deque()
        done = []
        i = 0
        t0 = time.perf_counter()
        while len(done) < len(reqs):
            now = time.perf_counter() - t0
            while i < len(reqs) and (not realtime or reqs[i].t_arrival <= now):
                queue.append(reqs[i])
                i += 1
            progressed = False
            for slot in self.slots:
                if slot.job is not None and slot.event.query():
                    req, t_start, host = slot.job
                    t_end = time.perf_counter() - t0
                    arrival = req.t_arrival if realtime else 0.0
                    if self.on_result is not None:
                        self.on_result(req, host)
                    self.store.release(req.key)
                    done.append(Completion(req.rid, t_end - arrival, t_start - arrival, int(np.prod(req.grid)),
                                           slot.index))
                    slot.job = None
                    progressed = True
            for slot in self.slots:
                if not queue:
                    break
                if slot.job is not None:
                    continue
                req = queue[0]
                plan = self.plan_of(req)
                if not self._fits(req, slot, plan):
                    if all(s.job is None for s in self.slots):
                        raise RuntimeError(f"job {req.rid} on grid {req.grid} does not fit the HBM budget "
                                           f"of {self.budget / 1e9:.1f} GB")
                    self.deferred += 1
                    break
                queue.popleft()
                self._start(req, slot, plan, t0)
                progressed = True
            if not progressed:
                idle = all(s.job is None for s in self.slots) and not queue
                if realtime and idle and i < len(reqs):
                    time.sleep(max(0.0, min(reqs[i].t_arrival - (time.perf_counter() - t0), 1e-3)))
                else:
                    time.sleep(poll_s)
        self.elapsed_s = time.perf_counter() - t0
        return done


def latency_summary(done, elapsed_s):
    lat = np.array([d.latency_s for d in done])
    q = np.array([d.queue_s for d in done])
    vox = sum(d.n for d in done)
    return {"requests": len(done), "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3),
            "queue_p99_ms": float(np.percentile(q, 99) * 1e3), "elapsed_s": elapsed_s,
            "volumes_per_s": len(done) / elapsed_s, "voxels_per_s": vox / elapsed_s}
