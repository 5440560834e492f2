"""Host logic of the batched and served paths (no GPU): stream assignment, the request
stream, the snapshot store's residency and eviction, and the workload generators' sizes."""

import numpy as np

from qpop_diffraction import workloads
from qpop_diffraction.batch import assign_streams
from qpop_diffraction.serve import SnapshotStore, latency_summary, poisson_requests, Completion
from qpop_diffraction._native import load_library


def test_assign_streams_balances_longest_first():
    costs = [9.0, 7.0, 6.0, 5.0, 4.0, 3.0, 2.0, 2.0]
    where, load = assign_streams(costs, 3)
    assert len(where) == len(costs)
    for s in range(3):
        assert np.isclose(load[s], sum(c for c, w in zip(costs, where) if w == s))
    assert max(load) - min(load) <= max(costs)
    assert sum(load) == sum(costs)
    # the longest job opens a stream of its own
    assert where[0] != where[1] and where[0] != where[2]


def test_poisson_requests_rate_and_draws():
    reqs = poisson_requests(4000, rate=20.0, grids=workloads.SERVE_GRIDS, n_snapshots=16, n_reflections=4, seed=1)
    t = np.array([r.t_arrival for r in reqs])
    assert np.all(np.diff(t) >= 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 1 / 20.0) < 0.1 / 20.0
    # exponential gaps: standard deviation equals the mean
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.1
    assert {r.grid for r in reqs} == set(workloads.SERVE_GRIDS)
    assert {r.snapshot for r in reqs} == set(range(16))
    assert {r.reflection for r in reqs} == set(range(4))
    again = poisson_requests(4000, rate=20.0, grids=workloads.SERVE_GRIDS, n_snapshots=16, n_reflections=4, seed=1)
    assert again == reqs
    burst = poisson_requests(5, rate=0.0, grids=workloads.SERVE_GRIDS, n_snapshots=2, n_reflections=4)
    assert all(r.t_arrival == 0.0 for r in burst)


class _Snap:
    def __init__(self, n):
        import torch
        self.P = torch.zeros((3, n), dtype=torch.float32)
        self.U = torch.zeros((3, n), dtype=torch.float64)
        self.phase = torch.zeros(n, dtype=torch.int8)


def test_snapshot_store_lru_and_pins():
    loads = []

    def spec(key):
        return workloads.snapshot_spec(*key[0])

    def load(key):
        loads.append(key)
        return _Snap(int(np.prod(key[0])))

    st = SnapshotStore(spec, load)
    g = (4, 4, 2)
    a, b, c = (g, 0), (g, 1), (g, 2)
    assert st.footprint(a) == spec(a).nbytes == 32 * 37
    st.acquire(a)
    st.acquire(b)
    st.release(b)
    st.acquire(c)
    st.release(c)
    assert st.nbytes == 3 * 32 * 37 and st.footprint(a) == 0
    st.acquire(b)  # b becomes most recently used
    st.release(b)
    # a is in use: the least recently used free one (c) goes first, then b, never a
    assert st.evict_one() == 32 * 37 and not st.resident(c)
    assert st.evict_one() == 32 * 37 and not st.resident(b)
    assert st.evict_one() == 0 and st.resident(a)
    st.release(a)
    assert st.evict_one(protect=a) == 0
    assert st.evict_one() == 32 * 37 and st.nbytes == 0
    st.acquire(c)
    assert loads == [a, b, c, c] and st.loads == 4 and st.evictions == 3


def test_latency_summary():
    done = [Completion(i, latency_s=(i + 1) * 1e-3, queue_s=0.0, n=10, slot=0) for i in range(100)]
    s = latency_summary(done, elapsed_s=2.0)
    assert np.isclose(s["p50_ms"], 50.5) and np.isclose(s["p99_ms"], 99.01)
    assert s["volumes_per_s"] == 50.0 and s["voxels_per_s"] == 500.0


def test_batch_grids_use_every_radix():
    lib = load_library()
    grids = workloads.batch_grids(48, seed=3)
    assert len(grids) == 48
    for L, d in grids:
        assert L in workloads.LATERALS and d in workloads.DEPTHS
    for n in workloads.LATERALS + workloads.DEPTHS:
        assert lib.qpd_fft_length_supported(n) == 1, n
    primes = set()
    for n in workloads.DEPTHS + workloads.LATERALS:
        m, p = n, 2
        while m > 1:
            while m % p == 0:
                primes.add(p)
                m //= p
            p += 1
    assert primes == {2, 3, 5, 7, 11, 13, 17}
