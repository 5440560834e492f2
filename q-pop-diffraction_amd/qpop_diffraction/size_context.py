"""Grid, phases, temperature and strain reference of a FerroDyn run.

Mirrors ``SizeContext`` of the reference (src/python/size_context.py:14-284):
the same attribute names, the same rules for the temperature ramp, lattice(T),
the average in-plane strain and the out-of-plane mismatch per material, and
the same crop semantics.  Lengths are in nm.
"""

import json
import math
from pathlib import Path

import numpy as np

from .constants import L0


class SizeContext:
    """Simulation box of one FerroDyn run (nm), with the diffraction phases."""

    def __init__(self):
        self.lx = self.ly = self.lz = 0.0
        self.nx = self.ny = self.nz = 0
        self.ns = self.nf = 0
        self.nPhase = 0
        self.nStruc = 0
        self.strainAvg = np.zeros(6)
        self.n = 0
        self.dx = self.dy = self.dz = 0.0
        self.k1 = self.k2 = 0
        self.kt = 0.0
        self.ns1 = self.k0 = 0
        self.hf = self.hs = self.h1 = self.h2 = self.hd = self.h = self.hstep = 0.0
        self.Rn1 = self.Rn2 = self.Rn3 = 0
        self.Cn1 = self.Cn2 = self.Cn3 = 0
        self.setting = {}
        self.names = []
        self.materials = []
        self.film = 0

    # -- system_setting.json (size_context.py:69-107) -------------------------
    def read_json(self, filename):
        """Read Mesh, materials and phases; return the material ids of the phases.

        A phase is a material whose ``Ferroelectric.Is_FE`` is true, in
        ``MaterialNames`` order.  Rotated runs raise ValueError.
        """
        path = Path(filename)
        self.setting = json.loads(path.read_text())
        if self.setting.get("Rotation", {}).get("Enabled", False):
            raise ValueError("Rotated runs (Rotation.Enabled) are not supported.")

        mesh = self.setting["Mesh"]
        self.nx, self.ny, self.nz = mesh["NX"], mesh["NY"], mesh["NZ"]
        self.lx = self.nx * mesh["DX"] / L0
        self.ly = self.ny * mesh["DY"] / L0
        self.lz = self.nz * mesh["DZ"] / L0
        self.ns = self.nf = 0
        self.nStruc = 3

        self.names = self.setting["Materials"]["MaterialNames"]
        mat_dir = path.parent / "materials"
        self.materials = [json.loads((mat_dir / f"{name}.json").read_text()) for name in self.names]
        fe_ids = [m for m, spec in enumerate(self.materials)
                  if spec.get("Ferroelectric", {}).get("Is_FE", False)]
        self.nPhase = len(fe_ids)
        self.film = fe_ids[0]

        print("\nRead parameters:")
        print(f"  lx = {self.lx}, ly = {self.ly}, lz = {self.lz}")
        print(f"  nx = {self.nx}, ny = {self.ny}, nz = {self.nz}")
        print(f"  nPhase = {self.nPhase} ({[self.names[m] for m in fe_ids]}), nStruc = {self.nStruc}\n")
        return fe_ids

    # -- temperature and lattice (size_context.py:109-122) ---------------------
    def temperature(self, step):
        """Snapshot temperature: FerroDyn's stepped ramp, clamped at Final."""
        spec = self.setting["SimulationControls"]["Temperature"]
        if not spec.get("Transient", False):
            return spec["Initial"]
        T = spec["Initial"] + spec["Step"] * (step // spec["Interval"])
        if spec["Step"] > 0:
            return min(T, spec["Final"])
        return max(T, spec["Final"])

    def lattice(self, m, T):
        """(a, b, c) of material m at temperature T, nm: a_i (1 + alpha_ii (T - T_ref))."""
        mech = self.materials[m]["Mechanical"]
        lc = mech["LatticeConstants"]
        alpha = mech["ThermalExpansion"]
        t_ref = lc["ReferenceTemperature"]
        return np.array([lc[key] * (1.0 + alpha[i][i] * (T - t_ref)) for i, key in enumerate("abc")]) / L0

    # -- strain reference (size_context.py:124-159) -----------------------------
    def reference_strain(self, aC, T):
        """Set strainAvg (film in-plane strain vs aC); return c_m(T)/aC_z - 1 per material."""
        init = self.setting["Solvers"].get("ElasticStrain", {}).get("InitialConditions", {})
        misfit = init.get("MisfitStrain", {})
        applied = init.get("AppliedStrain", {})
        layers = self.setting["Materials"]["LayerTypes"]
        air = {m for m, name in enumerate(self.names) if name in ("Air", "Vacuum")}

        # The bottom layer is a substrate when it is solid and some other solid layer exists.
        bottom = layers[0]
        other_solid = any(t != bottom and t not in air for t in layers)
        substrate = bottom if (bottom not in air and other_solid) else None

        a = self.lattice(self.film, T)
        self.strainAvg = np.zeros(6)
        geometry = self.setting["BoundaryConditions"].get("GeometryType", "Bulk").lower()
        if geometry == "thin-film":
            if misfit.get("Enabled", False):
                a[:2] = a[:2] * (1.0 + np.array(misfit["Strain"][:2]))
                self.strainAvg[5] = misfit["Strain"][2]
            elif substrate is not None:
                a[:2] = self.lattice(substrate, T)[:2]
        self.strainAvg[:2] = a[:2] / aC[:2] - 1.0
        if applied.get("Enabled", False) and str(self.film) in applied:
            self.strainAvg = self.strainAvg + applied[str(self.film)]

        print(f"  T = {T} K, strainAvg = {self.strainAvg}")
        c_of = np.array([self.lattice(m, T)[2] for m in range(len(self.names))])
        return c_of / aC[2] - 1.0

    # -- derived sizes (size_context.py:161-208) --------------------------------
    def setup(self):
        """Recompute n, spacings and the legacy film indices from the current box."""
        self.n = self.nx * self.ny * self.nz
        self.Rn1, self.Rn2, self.Rn3 = self.nx, self.ny, self.nz
        self.Cn1, self.Cn2, self.Cn3 = self.nx, self.ny, math.floor(self.nz / 2) + 1
        self.dx = self.lx / self.nx
        self.dy = self.ly / self.ny
        self.dz = self.lz / self.nz
        self.ns1 = 0
        self.k0 = self.ns1 + 1
        if self.nf == 0:
            self.k1, self.k2 = 1, self.nz
            self.hf = self.nf * self.dz
            self.hs = (self.ns - self.ns1 - 1) * self.dz
            self.h1 = self.dz * (self.k1 - 1)
            self.h2 = self.dz * (self.k2 - 1)
            self.hd = self.h2 - self.h1
            self.h = max(self.hs, self.hf)
            self.hstep = self.dz
        else:
            self.k1 = self.ns + 1
            self.k2 = min(self.ns + self.nf, self.nz)
        return True

    # -- crop (size_context.py:210-284) -----------------------------------------
    def apply_slice(self, x_limits, y_limits, z_limits):
        """Shrink the box to inclusive index limits (None = whole axis); return (x0,x1,y0,y1,z0,z1).

        Raises IndexError when an end index is outside the full grid.
        """
        spacing = (self.lx / self.nx, self.ly / self.ny, self.lz / self.nz)
        full = (self.nx, self.ny, self.nz)
        bounds = []
        for lim, n, axis in zip((x_limits, y_limits, z_limits), full, "xyz"):
            lo, hi = (lim[0], lim[1]) if lim is not None else (0, n - 1)
            if hi >= n:
                raise IndexError(f"{axis}_limits end {hi} is out of range for n{axis}={n}.")
            bounds.append((lo, hi))
        self.nx, self.ny, self.nz = (hi - lo + 1 for lo, hi in bounds)
        self.lx = self.nx * spacing[0]
        self.ly = self.ny * spacing[1]
        self.lz = self.nz * spacing[2]
        self.ns = self.nf = 0
        self.setup()
        (x0, x1), (y0, y1), (z0, z1) = bounds
        print(f"  Slice applied  : x=[{x0},{x1}], y=[{y0},{y1}], z=[{z0},{z1}]")
        print(f"  Slice grid     : nx={self.nx}, ny={self.ny}, nz={self.nz}")
        print(f"  Slice domain   : lx={self.lx:.4f} nm, ly={self.ly:.4f} nm, lz={self.lz:.4f} nm")
        return x0, x1, y0, y1, z0, z1
