"""Device-side generators of the benchmark workloads (PyTorch on the GPU).

Each generator returns a ``Workload``: the run's SizeContext and AtomList and
device tensors in FerroDyn order (x fastest, memory dims (nx, ny, nz)):
P (3, n) float32/64 in C/m^2, U (3, n) float64 in nm with the u_z reference
ramp already added (as cli.py:202-217 of the reference adds it), phase (n,) int8.

  * two_phase:      periodic Voronoi grains, each tetragonal (P = +-0.6 z, a second
                    ferroelectric material with c/a = 1.06, entering through the
                    u_z ramp) or rhombohedral (P = 0.4 (+-1, +-1, +-1)/sqrt 3);
                    u = 0.5 % in-plane ramp plus a Gaussian random field of rms
                    0.005 nm; reflection 002 (512 x 512 x 104 in the benchmark);
  * twin_stack:     two-variant 90-degree twins along x, period 8-32 cells,
                    |P| 0.5-0.7 C/m^2;
  * hex_twin:       three in-plane variants P = 0.6 (cos t, sin t, 0), t = 0/120/240 deg.
"""

from __future__ import annotations

import math
import os
import tempfile
from dataclasses import dataclass

import numpy as np

from .atom_list import AtomList
from .size_context import SizeContext
from . import synth


@dataclass
class Workload:
    params: SizeContext
    atoms: AtomList
    dims: tuple
    axis: tuple
    P: object
    U: object
    phase: object
    name: str


def atoms_from_text(text, n_phase):
    fd, path = tempfile.mkstemp(suffix=".atom.in")
    try:
        with os.fdopen(fd, "w") as fh:
            fh.write(text)
        atoms = AtomList()
        assert atoms.read(path, n_phase, 3) and atoms.setup()
    finally:
        os.remove(path)
    return atoms


def box(nx, ny, nz, h, n_phase):
    p = SizeContext()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.lx, p.ly, p.lz = nx * h, ny * h, nz * h
    p.nPhase, p.nStruc = n_phase, 3
    p.strainAvg = np.zeros(6)
    p.setup()
    return p


def _grid(nx, ny, nz, device):
    import torch
    z = torch.arange(nz, device=device, dtype=torch.float32).view(nz, 1, 1)
    y = torch.arange(ny, device=device, dtype=torch.float32).view(1, ny, 1)
    x = torch.arange(nx, device=device, dtype=torch.float32).view(1, 1, nx)
    return z, y, x


def grf(shape, corr, rms, gen, device):
    """Periodic Gaussian random field (z, y, x): white noise smoothed by a separable Gaussian."""
    import torch
    f = torch.randn(shape, generator=gen, device=device, dtype=torch.float64)
    for axis, n in enumerate(shape):
        if n < 2:
            continue
        d = np.minimum(np.arange(n), n - np.arange(n)).astype(np.float64)
        kern = np.exp(-0.5 * (d / corr) ** 2)
        kern /= kern.sum()
        out = torch.zeros_like(f)
        for s in np.nonzero(kern > 1e-10)[0]:
            out += float(kern[s]) * torch.roll(f, int(s), dims=axis)
        f = out
    return f * (rms / f.pow(2).mean().sqrt())


def voronoi(shape, n_grains, gen, device):
    import torch
    nz, ny, nx = shape
    seeds = torch.rand((n_grains, 3), generator=gen, device=device, dtype=torch.float32) * torch.tensor(
        [nz, ny, nx], device=device, dtype=torch.float32)
    z, y, x = _grid(nx, ny, nz, device)
    best = torch.full(shape, float("inf"), device=device)
    gid = torch.zeros(shape, dtype=torch.int64, device=device)
    for g in range(n_grains):
        dz = (z + 0.5 - seeds[g, 0]).abs(); dz = torch.minimum(dz, nz - dz)
        dy = (y + 0.5 - seeds[g, 1]).abs(); dy = torch.minimum(dy, ny - dy)
        dx = (x + 0.5 - seeds[g, 2]).abs(); dx = torch.minimum(dx, nx - dx)
        d = dz * dz + dy * dy + dx * dx
        closer = d < best
        best = torch.where(closer, d, best)
        gid = torch.where(closer, torch.full_like(gid, g), gid)
    return gid


def uz_ramp(dc_of_cell, dz):
    """(cumsum_z dc - dc/2) dz over the full stack (cli.py:207-208), dc_of_cell shaped (z, y, x)."""
    import torch
    return (torch.cumsum(dc_of_cell, dim=0) - 0.5 * dc_of_cell) * dz


def two_phase(nx=512, ny=512, nz=104, n_grains=40, seed=2, a=synth.A_PTO, precision=32, device="cuda"):
    import torch
    gen = torch.Generator(device=device).manual_seed(seed)
    shape = (nz, ny, nx)
    gid = voronoi(shape, n_grains, gen, device)
    tetra = torch.rand(n_grains, generator=gen, device=device) < 0.5
    signs = torch.where(torch.rand((n_grains, 3), generator=gen, device=device) < 0.5, -1.0, 1.0)
    is_t = tetra[gid]
    s = signs[gid]  # (z, y, x, 3)
    P = torch.empty((3,) + shape, device=device, dtype=torch.float64)
    r = 0.4 / math.sqrt(3.0)
    P[0] = torch.where(is_t, 0.0, r * s[..., 0])
    P[1] = torch.where(is_t, 0.0, r * s[..., 1])
    P[2] = torch.where(is_t, 0.6 * s[..., 2], r * s[..., 2])
    z, y, x = _grid(nx, ny, nz, device)
    U = torch.stack([grf(shape, 4.0, 0.005, gen, device) for _ in range(3)])
    U[0] += (0.005 * a) * x.double()
    U[1] += (0.005 * a) * y.double()
    # materials: 0 tetragonal (c = 1.06 a), 1 rhombohedral (c = a); aC = a: dc = 0.06 / 0
    dc = torch.where(is_t, 0.06, 0.0).double()
    U[2] += uz_ramp(dc, a)
    phase = torch.where(is_t, 0, 1).to(torch.int8)
    params = box(nx, ny, nz, a, 2)
    atoms = atoms_from_text(synth.atom_file(synth.q00_of((0, 0, 2), a), a=a, n_phase=2), 2)
    real = torch.float32 if precision == 32 else torch.float64
    return Workload(params, atoms, (nx, ny, nz), (0, 1, 2), P.to(real).reshape(3, -1).contiguous(),
                    U.reshape(3, -1).contiguous(), phase.reshape(-1).contiguous(),
                    f"two-phase Voronoi {nx}x{ny}x{nz}, {n_grains} grains")


def twin_stack(nx, ny, nz, seed, hkl=(0, 0, 2), a=synth.A_PTO, precision=32, device="cuda"):
    import torch
    rng = np.random.default_rng(seed)
    period = int(rng.integers(8, 33))
    mag = float(rng.uniform(0.5, 0.7))
    variant = ((torch.arange(nx, device=device) // max(1, period // 2)) % 2).view(1, 1, nx).expand(nz, ny, nx)
    P = torch.zeros((3, nz, ny, nx), device=device, dtype=torch.float64)
    P[0] = mag * (variant == 0)
    P[2] = mag * (variant == 1)
    U = torch.zeros_like(P)
    phase = torch.zeros(nx * ny * nz, dtype=torch.int8, device=device)
    params = box(nx, ny, nz, a, 1)
    atoms = atoms_from_text(synth.atom_file(synth.q00_of(hkl, a), a=a), 1)
    real = torch.float32 if precision == 32 else torch.float64
    return Workload(params, atoms, (nx, ny, nz), (0, 1, 2), P.to(real).reshape(3, -1).contiguous(),
                    U.reshape(3, -1).contiguous(), phase, f"twin stack {nx}x{ny}x{nz} period {period}")


def hex_twin(nx, ny, nz, seed, hkl=(0, 0, 2), a=synth.A_PTO, precision=32, device="cuda"):
    import torch
    gen = torch.Generator(device=device).manual_seed(seed)
    gid = voronoi((nz, ny, nx), 12, gen, device) % 3
    theta = torch.deg2rad(torch.tensor([0.0, 120.0, 240.0], device=device, dtype=torch.float64))[gid]
    P = torch.stack([0.6 * torch.cos(theta), 0.6 * torch.sin(theta), torch.zeros_like(theta)])
    U = torch.zeros_like(P)
    phase = torch.zeros(nx * ny * nz, dtype=torch.int8, device=device)
    params = box(nx, ny, nz, a, 1)
    atoms = atoms_from_text(synth.atom_file(synth.q00_of(hkl, a), a=a), 1)
    real = torch.float32 if precision == 32 else torch.float64
    return Workload(params, atoms, (nx, ny, nz), (0, 1, 2), P.to(real).reshape(3, -1).contiguous(),
                    U.reshape(3, -1).contiguous(), phase, f"hex twin {nx}x{ny}x{nz}")


def host_inputs(w):
    """The workload in the reference diffraction() layout: (x, y, z) C order, float64 numpy."""
    nx, ny, nz = w.dims
    P = w.P.double().reshape(3, nz, ny, nx).permute(0, 3, 2, 1).contiguous().cpu().numpy().ravel()
    U = w.U.reshape(3, nz, ny, nx).permute(0, 3, 2, 1).contiguous().cpu().numpy().ravel()
    ph = w.phase.reshape(nz, ny, nx).permute(2, 1, 0).contiguous().cpu().numpy()
    n_phase = w.params.nPhase
    oPhase = np.stack([(ph == p).astype(np.float64) for p in range(n_phase)]).ravel()
    region = (ph >= 0).astype(np.float64).ravel()
    return oPhase, P, U, region


# -- snapshots and reflections of the batched and served workloads ------------

REFLECTIONS = ((0, 0, 2), (1, 0, 3), (2, 0, 2), (1, 1, 3))
LATERALS = (128, 192, 256, 320, 384, 448, 512)
# depths 24-104 chosen so that every radix of the transforms appears: 2, 3, 4, 5, 7, 8, 11, 13, 17
DEPTHS = (24, 30, 34, 40, 44, 52, 56, 68, 84, 88, 104)


def pto_atoms(hkl, n_phase=1, a=synth.A_PTO):
    """The 5-atom PbTiO3-like cell with the pattern centred on reflection hkl."""
    return atoms_from_text(synth.atom_file(synth.q00_of(hkl, a), a=a, n_phase=n_phase), n_phase)


@dataclass
class Snapshot:
    """Device-resident fields of one snapshot (FerroDyn order, x fastest)."""
    key: tuple
    params: SizeContext
    dims: tuple
    axis: tuple
    P: object
    U: object
    phase: object

    def nbytes(self):
        return sum(t.numel() * t.element_size() for t in (self.P, self.U, self.phase))


def twin_snapshot(key, nx, ny, nz, seed, precision=32, a=synth.A_PTO, device="cuda"):
    """Two-variant 90-degree twin stack along x: P || x <-> P || z, period 8-32 cells, |P| 0.5-0.7."""
    import torch
    rng = np.random.default_rng(seed)
    period = int(rng.integers(8, 33))
    mag = float(rng.uniform(0.5, 0.7))
    variant = ((torch.arange(nx, device=device) // max(1, period // 2)) % 2).view(1, 1, nx).expand(nz, ny, nx)
    gen = torch.Generator(device=device).manual_seed(seed)
    P = 0.01 * torch.randn((3, nz, ny, nx), generator=gen, device=device, dtype=torch.float64)
    P[0] += mag * (variant == 0)
    P[2] += mag * (variant == 1)
    real = torch.float32 if precision == 32 else torch.float64
    n = nx * ny * nz
    return Snapshot(key, box(nx, ny, nz, a, 1), (nx, ny, nz), (0, 1, 2), P.to(real).reshape(3, n).contiguous(),
                    torch.zeros((3, n), dtype=torch.float64, device=device),
                    torch.zeros(n, dtype=torch.int8, device=device))


def hex_snapshot(key, nx, ny, nz, seed, precision=32, a=synth.A_PTO, device="cuda"):
    """Three in-plane variants P = 0.6 (cos t, sin t, 0), t = 0/120/240 degrees, in periodic Voronoi domains."""
    import torch
    gen = torch.Generator(device=device).manual_seed(seed)
    gid = voronoi((nz, ny, nx), 12, gen, device) % 3
    theta = torch.deg2rad(torch.tensor([0.0, 120.0, 240.0], device=device, dtype=torch.float64))[gid]
    P = torch.stack([0.6 * torch.cos(theta), 0.6 * torch.sin(theta), torch.zeros_like(theta)])
    real = torch.float32 if precision == 32 else torch.float64
    n = nx * ny * nz
    return Snapshot(key, box(nx, ny, nz, a, 1), (nx, ny, nz), (0, 1, 2), P.to(real).reshape(3, n).contiguous(),
                    torch.zeros((3, n), dtype=torch.float64, device=device),
                    torch.zeros(n, dtype=torch.int8, device=device))


def batch_grids(n_snapshots=48, seed=3):
    """Seeded (lateral, depth) of each snapshot of the batched workload."""
    rng = np.random.default_rng(seed)
    return [(int(rng.choice(LATERALS)), int(rng.choice(DEPTHS))) for _ in range(n_snapshots)]


@dataclass
class SnapshotSpec:
    """What a plan and the HBM budget need to know of a snapshot before its fields exist."""
    params: SizeContext
    dims: tuple
    axis: tuple
    nbytes: int


def snapshot_spec(nx, ny, nz, precision=32, a=synth.A_PTO):
    n = nx * ny * nz
    real = 4 if precision == 32 else 8
    return SnapshotSpec(box(nx, ny, nz, a, 1), (nx, ny, nz), (0, 1, 2), n * (3 * real + 3 * 8 + 1))


SERVE_GRIDS = ((128, 128, 32), (256, 256, 64), (512, 512, 104))


def serve_store(precision=32, seed0=1000):
    """SnapshotStore of the served workload: key (grid, id) -> hexagonal-twin snapshot of seed seed0 + id."""
    from .serve import SnapshotStore
    specs = {}

    def spec(key):
        if key[0] not in specs:
            specs[key[0]] = snapshot_spec(*key[0], precision=precision)
        return specs[key[0]]

    def load(key):
        grid, sid = key
        return hex_snapshot(key, *grid, seed=seed0 + int(sid), precision=precision)

    return SnapshotStore(spec, load)


# -- the large single volume -----------------------------------------------------

A_STO = 0.3905  # nm, SrTiO3-like substrate


@dataclass
class LargeVolume:
    """Device fields of one volume in the reference's (x, y, z) C order (z fastest)."""
    params: SizeContext
    atoms: AtomList
    dims: tuple   # memory dims, fastest first: (nz, ny, nx)
    axis: tuple   # (2, 1, 0)
    P: object
    U: object
    phase: object
    name: str


def vortex(nx=4000, ny=4000, nz=68, period_x=10, period_z=12, substrate=16, noise=0.05, strain=-0.01, seed=5,
           precision=32, a=synth.A_PTO, slab=200, device="cuda"):
    """Polar-vortex superlattice on a substrate, generated in x slabs straight into the z-fastest layout.

    Film (z >= substrate): P = 0.5 (sin(2 pi z / period_z) cos(2 pi x / period_x), 0, cos(2 pi z / period_z))
    C/m^2 plus N(0, noise * 0.5) on each component; substrate cells (z < substrate) have P = 0 and
    phase -1 (a non-ferroelectric layer: no scattering, but its lattice mismatch enters the
    u_z ramp as dc = a_STO / a - 1).  u = strain * (x dx, y dy) in-plane plus that ramp.
    """
    import torch
    n = nx * ny * nz
    real = torch.float32 if precision == 32 else torch.float64
    P = torch.empty((3, n), dtype=real, device=device)
    U = torch.empty((3, n), dtype=torch.float64, device=device)
    phase = torch.empty(n, dtype=torch.int8, device=device)
    gen = torch.Generator(device=device).manual_seed(seed)
    z = torch.arange(nz, device=device, dtype=torch.float64)
    y = torch.arange(ny, device=device, dtype=torch.float64)
    dc = torch.where(z < substrate, A_STO / a - 1.0, 0.0)
    uz = (torch.cumsum(dc, 0) - 0.5 * dc) * a  # (nz,)
    film = (z >= substrate).view(1, 1, nz)
    amp = 0.5
    for x0 in range(0, nx, slab):
        x1 = min(nx, x0 + slab)
        x = torch.arange(x0, x1, device=device, dtype=torch.float64).view(-1, 1, 1)
        shape = (x1 - x0, ny, nz)
        sz = torch.sin(2 * math.pi * z / period_z).view(1, 1, nz)
        cz = torch.cos(2 * math.pi * z / period_z).view(1, 1, nz)
        cx = torch.cos(2 * math.pi * x / period_x)
        lo, hi = x0 * ny * nz, x1 * ny * nz
        comps = (amp * sz * cx, torch.zeros((1, 1, 1), device=device, dtype=torch.float64), amp * cz)
        for d in range(3):
            val = comps[d].expand(shape) + (noise * amp) * torch.randn(shape, generator=gen, device=device,
                                                                          dtype=torch.float64)
            P[d, lo:hi] = torch.where(film, val, 0.0).reshape(-1).to(real)
        U[0, lo:hi] = (strain * a * x).expand(shape).reshape(-1)
        U[1, lo:hi] = (strain * a * y.view(1, ny, 1)).expand(shape).reshape(-1)
        U[2, lo:hi] = uz.view(1, 1, nz).expand(shape).reshape(-1)
        phase[lo:hi] = torch.where(film, 0, -1).to(torch.int8).expand(shape).reshape(-1)
    params = box(nx, ny, nz, a, 1)
    params.strainAvg = np.array([strain, strain, 0.0, 0.0, 0.0, 0.0])
    atoms = pto_atoms((0, 0, 2))
    return LargeVolume(params, atoms, (nz, ny, nx), (2, 1, 0), P, U, phase,
                       f"polar-vortex superlattice {nx}x{ny}x{nz} on a {substrate}-cell substrate")
