"""Kinematic diffraction on one MI355X (reference: src/python/diffraction.py:1-113).

For each atom a of the unit cell, with u_a(r) = u(r) + su_a . P(r) in nm:

    fE_a(r) = sum_phases f_a w(r) exp(-i qC . u_a(r))
    A(k)    = S(k) sum_a exp(-i (k + G) . R_a) / V [FFT(fE_a)/n - i k . FFT(fE_a u_a)/n]
    I       = fftshift(max(|A|^2, 1e-50)),   q = qC + k.

``reciprocal_centre``, ``wavevectors`` and ``q_axes`` are the reference's FP64
host formulas.  ``diffraction`` has the reference's signature and return value
and runs on the GPU through libqpd (three HIP passes per volume).  The working
precision is complex FP32 with the phase reduced in FP64 (default), or FP64
throughout; pick it with ``precision=`` or ``QPOP_DIFFRACTION_PRECISION``.
There is no CPU fallback: without the library or a GPU the call raises.
"""

from __future__ import annotations

import math
import os

import numpy as np

from ._native import Plan


def reciprocal_centre(params, atoms):
    """(G, qC): the reciprocal lattice point nearest q00 and G - eps.G, nm^-1 (diffraction.py:36-48)."""
    aC = atoms.aC
    hkl = np.round(aC @ atoms.q00 / (2 * math.pi))
    G = hkl @ (2 * math.pi * np.linalg.inv(aC).T)
    s = params.strainAvg
    eps = np.array([[s[0], s[5], s[4]], [s[5], s[1], s[3]], [s[4], s[3], s[2]]])
    qC = G - eps @ G
    print(f"  G = {G}, qC = {qC}")
    return G, qC


def wavevectors(params):
    """FFT wavevectors along x, y, z in nm^-1, FFT order (diffraction.py:51-54)."""
    shape = (params.nx, params.ny, params.nz)
    return [2 * math.pi * np.fft.fftfreq(m, d) for m, d in zip(shape, (params.dx, params.dy, params.dz))]


def q_axes(params, qC):
    """1-D q axes of the fftshifted intensity, nm^-1 (diffraction.py:57-59)."""
    return tuple(np.fft.fftshift(kd) + qC[d] for d, kd in enumerate(wavevectors(params)))


def default_precision():
    p = int(os.environ.get("QPOP_DIFFRACTION_PRECISION", "32"))
    if p not in (32, 64):
        raise ValueError("QPOP_DIFFRACTION_PRECISION must be 32 or 64")
    return p


def crystal_tables(params, atoms):
    """Everything a plan needs from the crystal: R, f, su stack, G, qC, V (diffraction.py:78-94)."""
    G, qC = reciprocal_centre(params, atoms)
    V = float(np.linalg.det(atoms.aC))
    R = atoms.xAtom @ atoms.aC
    su = np.stack([atoms.su1, atoms.su2, atoms.su3])  # (direction, nStruc, nPhase, nAtom)
    return R, atoms.fAtom, su, G, qC, V


class Volume:
    """Device-resident inputs of one volume in a given memory layout.

    dims: memory extents, dims[0] fastest; axis: physical axis of each dim.
    P (3, n) and U (3, n) hold physical components x, y, z; phase (n,) int8 is the
    phase index of a cell (-1: does not scatter); fraction (nPhase, n) and region
    (n,) are optional weights.  Arrays are host numpy or device tensors.
    """

    def __init__(self, dims, axis, P, U, phase, fraction=None, region=None, precision=32, device="cuda"):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("qpop_diffraction needs a ROCm GPU (torch.cuda.is_available() is False)")
        real = torch.float32 if precision == 32 else torch.float64
        n = int(np.prod(dims))

        def dev(a, dtype, shape):
            t = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a))
            t = t.to(device=device, dtype=dtype).reshape(shape).contiguous()
            return t

        self.dims, self.axis, self.precision, self.n = tuple(dims), tuple(axis), precision, n
        self.P = dev(P, real, (3, n))
        self.U = dev(U, torch.float64, (3, n))
        self.phase = dev(phase, torch.int8, (n,))
        self.fraction = None if fraction is None else dev(fraction, real, (-1, n))
        self.region = None if region is None else dev(region, real, (n,))

    def nbytes(self):
        ts = [self.P, self.U, self.phase, self.fraction, self.region]
        return sum(t.numel() * t.element_size() for t in ts if t is not None)


def weights_from_phases(oPhase, region, nPhase, n):
    """(phase index, fraction, region) equivalent to w_p = oPhase_p * region (diffraction.py:91).

    One-hot 0/1 phases become a phase index with an optional region; anything
    else becomes per-phase fractions (already multiplied by region).
    """
    w = np.asarray(oPhase, dtype=np.float64).reshape(nPhase, n)
    reg = np.asarray(region, dtype=np.float64).reshape(n)
    binary = np.all((w == 0.0) | (w == 1.0)) and np.all(w.sum(axis=0) <= 1.0)
    if binary:
        phase = np.where(w.any(axis=0), np.argmax(w, axis=0), -1).astype(np.int8)
        reg_out = None if np.all(reg[phase >= 0] == 1.0) else reg
        return phase, None, reg_out
    frac = w * reg[None, :]
    phase = np.where(frac.any(axis=0), 0, -1).astype(np.int8)
    return phase, frac, None


class Engine:
    """Plans and workspaces for one crystal on one GPU, cached by layout and precision.

    Workspaces are cached per stream: calls enqueued on different streams never share one.
    order and arith are the plan options of ``Plan`` (pass order, arithmetic of FP32 plans).
    """

    def __init__(self, params, atoms, order="auto", arith="auto"):
        self.params = params
        self.atoms = atoms
        self.tables = crystal_tables(params, atoms)
        self.order, self.arith = order, arith
        self._plans = {}
        self._work = {}

    @classmethod
    def from_tables(cls, params, tables, order="auto", arith="auto"):
        """An engine for crystal tables (R, f, su, G, qC, V) that were prepared elsewhere."""
        eng = cls.__new__(cls)
        eng.params, eng.atoms, eng.tables, eng._plans, eng._work = params, None, tuple(tables), {}, {}
        eng.order, eng.arith = order, arith
        return eng

    def plan(self, dims, axis, precision):
        key = (tuple(dims), tuple(axis), precision)
        if key not in self._plans:
            R, f, su, G, qC, V = self.tables
            h = (self.params.dx, self.params.dy, self.params.dz)
            self._plans[key] = Plan(dims, axis, h, R, f, su, G, qC, V, precision, order=self.order,
                                    arith=self.arith if precision == 32 else "auto")
        return self._plans[key]

    def workspace(self, plan, stream=None):
        """A workspace of the plan's size for calls on ``stream`` (None: torch's current stream)."""
        import torch
        sid = stream.value if stream is not None else torch.cuda.current_stream().cuda_stream
        key = (plan.workspace_bytes, sid)
        if key not in self._work:
            self._work[key] = torch.empty(max(1, plan.workspace_bytes), dtype=torch.uint8, device="cuda")
        return self._work[key]

    def tune(self, vol: Volume, cache=None, out=None):
        """Autotune the launch schedules of the plan for this volume's layout and precision."""
        import torch
        from .autotune import tune
        plan = self.plan(vol.dims, vol.axis, vol.precision)
        if out is None:
            out = torch.empty(vol.n, dtype=torch.float64, device="cuda")
        return tune(plan, vol.P, vol.U, vol.phase, out, self.workspace(plan), vol.fraction, vol.region, cache=cache)

    def intensity(self, vol: Volume, out=None, stream=None):
        """fftshifted intensity (device float64 tensor shaped like the layout) of a volume."""
        import torch
        plan = self.plan(vol.dims, vol.axis, vol.precision)
        if out is None:
            out = torch.empty(vol.n, dtype=torch.float64, device="cuda")
        plan.intensity(vol.P, vol.U, vol.phase, out, self.workspace(plan, stream), vol.fraction, vol.region, stream)
        return out


def diffraction(params, atoms, oPhase, oStruc, u, region, precision=None):
    """Drop-in for the reference ``diffraction`` (diffraction.py:62-113).

    Args are the reference's: oPhase (nPhase*n,), oStruc (3n,) P in C/m^2, u (3n,)
    in nm, region (n,), all in (x, y, z) C order.  Returns (I, qx, qy, qz) with
    I float64 (nx, ny, nz), fftshifted.
    """
    precision = default_precision() if precision is None else int(precision)
    nx, ny, nz = params.nx, params.ny, params.nz
    n = params.n
    # (x, y, z) C order is z fastest: memory dims (nz, ny, nx), physical axes (2, 1, 0)
    dims, axis = (nz, ny, nx), (2, 1, 0)
    phase, frac, reg = weights_from_phases(oPhase, region, params.nPhase, n)
    engine = Engine(params, atoms)
    vol = Volume(dims, axis, np.asarray(oStruc).reshape(params.nStruc, n), np.asarray(u).reshape(3, n),
                 phase, frac, reg, precision)
    I = engine.intensity(vol).cpu().numpy().reshape(nx, ny, nz)
    _, qC = engine.tables[3], engine.tables[4]
    return (I,) + q_axes(params, qC)
