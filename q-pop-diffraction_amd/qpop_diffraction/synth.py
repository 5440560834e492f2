"""Seeded synthetic FerroDyn runs for tests, fixtures and benchmarks.

Dependency-free apart from NumPy (it is also loaded by file path from the
fixture generator, which runs under another interpreter), so it uses no
relative imports.  Fields are returned in FerroDyn file order (component, z, y, x)
and SI units (P in C/m^2, u in m), as Q-POP-FerroDyn stores them.

Workloads (sizes nx x ny x nz):
  * perovskite_run:   PbTiO3-like 5-atom cell, two 180-degree z domains with tanh
                      walls plus noise, Gaussian-random-field u (64 x 64 x 32 grid);
  * two_phase_run:    periodic Voronoi grains, tetragonal (P = +-0.6 z, c/a 1.06) or
                      rhombohedral (P = 0.4 (+-1, +-1, +-1)/sqrt 3), in-plane strain
                      ramp plus GRF u (512 x 512 x 104 grid);
  * twin_run:         two-variant 90-degree twin stack along x (P || x <-> P || z);
  * hex_twin_run:     three-variant in-plane 120-degree domains.
"""

import json
import math
import os

import numpy as np

A_PTO = 0.39  # nm
FORM_PTO = (82.0, 22.0, 8.0, 8.0, 8.0)
POS_PTO = ((0.0, 0.0, 0.0), (0.5, 0.5, 0.5), (0.5, 0.5, 0.0), (0.5, 0.0, 0.5), (0.0, 0.5, 0.5))
SIGN_PTO = (1.0, -1.0, -1.0, -1.0, -1.0)
SHIFT_PER_P = 1e-11  # m per C/m^2 (0.01 nm per C/m^2)


def atom_file(q00, a=A_PTO, c=None, form=FORM_PTO, pos=POS_PTO, signs=SIGN_PTO, shift=SHIFT_PER_P, n_phase=1,
              aC=None, shifts=None):
    """parameter.atom.in text of a perovskite cell; every phase block uses the same crystal.

    The P -> shift coefficient of atom a is diagonal: signs[a] * shift along each
    axis, or shifts[a][d] (m per C/m^2) along axis d when ``shifts`` is given.
    """
    c = a if c is None else c
    aC = np.diag([a, a, c]) if aC is None else np.asarray(aC)
    if shifts is None:
        shifts = [[s * shift] * 3 for s in signs]
    num = lambda v: repr(float(v))  # plain decimal text whatever the scalar type
    lines = ["! synthetic perovskite cell", " ".join(num(v) for v in q00)]
    lines += [" ".join(num(v) for v in row) for row in aC]
    lines += [str(len(pos))]
    lines += [" ".join(num(v) for v in p) for p in pos]
    for ph in range(n_phase):
        lines.append(str(ph + 1))
        for f in form:
            fc = complex(f)
            lines.append(f"{num(fc.real)} {num(fc.imag)}")
        for sa in shifts:
            for k in range(3):
                lines.append(" ".join(repr(float(sa[d]) if d == k else 0.0) for d in range(3)))
    return "\n".join(lines) + "\n"


def q00_of(hkl, a=A_PTO, c=None):
    c = a if c is None else c
    return (2 * math.pi * hkl[0] / a, 2 * math.pi * hkl[1] / a, 2 * math.pi * hkl[2] / c)


def material(name, is_fe, a, b=None, c=None):
    return {
        "Name": name,
        "Ferroelectric": {"Is_FE": bool(is_fe)},
        "Mechanical": {
            "LatticeConstants": {"a": a, "b": a if b is None else b, "c": a if c is None else c,
                                 "ReferenceTemperature": 300.0},
            "ThermalExpansion": [[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]],
        },
    }


def gaussian_random_field(shape, corr, rms, rng):
    """Periodic GRF: white noise smoothed by a separable Gaussian of width corr cells, scaled to rms."""
    field = rng.standard_normal(shape)
    for axis, n in enumerate(shape):
        if n < 2:
            continue
        d = np.minimum(np.arange(n), n - np.arange(n)).astype(np.float64)
        kern = np.exp(-0.5 * (d / corr) ** 2)
        kern /= kern.sum()
        # circular convolution along the axis as a sum of rolls (deterministic, no FFT)
        out = np.zeros_like(field)
        for s in range(n):
            if kern[s] > 1e-12:
                out += kern[s] * np.roll(field, s, axis=axis)
        field = out
    std = np.sqrt(np.mean(field ** 2))
    return field * (rms / std) if std > 0 else field


def perovskite_run(nx=64, ny=64, nz=32, seed=1, a=A_PTO):
    """Two z domains (+-0.6 C/m^2) with tanh walls (width 2 cells) at x = nx/4, 3nx/4, noise 0.02, GRF u."""
    rng = np.random.default_rng(seed)
    x = np.arange(nx) + 0.5
    w = 2.0
    prof = np.tanh((x - nx / 4) / w) * np.tanh((3 * nx / 4 - x) / w)
    P = np.zeros((3, nz, ny, nx))
    P[2] = 0.6 * prof[None, None, :]
    P += rng.normal(0.0, 0.02, P.shape)
    u = np.stack([gaussian_random_field((nz, ny, nx), 4.0, 0.005, rng) for _ in range(3)]) * 1e-9
    mats = [material("Film", True, a * 1e-9)]
    mat = np.zeros((nz, ny, nx), dtype=np.int32)
    return {"P": P, "u": u, "mat": mat, "materials": mats, "spacing": a * 1e-9,
            "atom": atom_file(q00_of((0, 0, 2), a), a=a)}


def periodic_voronoi(shape, n_grains, rng):
    """Grain id of every cell of a periodic box (z, y, x) for n_grains random seeds."""
    nz, ny, nx = shape
    seeds = rng.random((n_grains, 3)) * np.array([nz, ny, nx])
    zz, yy, xx = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    best = np.full(shape, np.inf)
    gid = np.zeros(shape, dtype=np.int32)
    for g, (sz, sy, sx) in enumerate(seeds):
        dz = np.abs(zz - sz); dz = np.minimum(dz, nz - dz)
        dy = np.abs(yy - sy); dy = np.minimum(dy, ny - dy)
        dx = np.abs(xx - sx); dx = np.minimum(dx, nx - dx)
        d = dz * dz + dy * dy + dx * dx
        closer = d < best
        best[closer] = d[closer]
        gid[closer] = g
    return gid


def two_phase_run(nx=512, ny=512, nz=104, n_grains=40, seed=2, a=A_PTO):
    """Voronoi grains, each tetragonal (material 0, c = 1.06 a) or rhombohedral (material 1)."""
    rng = np.random.default_rng(seed)
    gid = periodic_voronoi((nz, ny, nx), n_grains, rng)
    tetra = rng.random(n_grains) < 0.5
    signs = rng.choice([-1.0, 1.0], size=(n_grains, 3))
    P = np.zeros((3, nz, ny, nx))
    for g in range(n_grains):
        cells = gid == g
        if tetra[g]:
            P[2][cells] = 0.6 * signs[g, 2]
        else:
            for d in range(3):
                P[d][cells] = 0.4 * signs[g, d] / math.sqrt(3.0)
    mat = np.where(tetra[gid], 0, 1).astype(np.int32)
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    u = np.stack([gaussian_random_field((nz, ny, nx), 4.0, 0.005, rng) for _ in range(3)]) * 1e-9
    u[0] += 0.005 * xx * a * 1e-9
    u[1] += 0.005 * yy * a * 1e-9
    mats = [material("Tetragonal", True, a * 1e-9, c=1.06 * a * 1e-9), material("Rhombohedral", True, a * 1e-9)]
    return {"P": P, "u": u, "mat": mat, "materials": mats, "spacing": a * 1e-9,
            "atom": atom_file(q00_of((0, 0, 2), a), a=a, n_phase=2)}


def twin_run(nx, ny, nz, seed, hkl=(0, 0, 2), a=A_PTO):
    """90-degree twin stack along x: P || x and P || z alternate with a random period of 8-32 cells."""
    rng = np.random.default_rng(seed)
    period = int(rng.integers(8, 33))
    mag = float(rng.uniform(0.5, 0.7))
    variant = ((np.arange(nx) // max(1, period // 2)) % 2)[None, None, :]
    P = np.zeros((3, nz, ny, nx))
    P[0] = mag * (variant == 0)
    P[2] = mag * (variant == 1)
    P = P + rng.normal(0.0, 0.01, P.shape)
    u = np.zeros((3, nz, ny, nx))
    mats = [material("Film", True, a * 1e-9)]
    mat = np.zeros((nz, ny, nx), dtype=np.int32)
    return {"P": P, "u": u, "mat": mat, "materials": mats, "spacing": a * 1e-9,
            "atom": atom_file(q00_of(hkl, a), a=a), "period": period}


def hex_twin_run(nx, ny, nz, seed, hkl=(0, 0, 2), a=A_PTO):
    """Three in-plane variants P = 0.6 (cos t, sin t, 0), t in {0, 120, 240} degrees, in Voronoi domains."""
    rng = np.random.default_rng(seed)
    gid = periodic_voronoi((nz, ny, nx), 12, rng) % 3
    theta = np.deg2rad(np.array([0.0, 120.0, 240.0]))[gid]
    P = np.stack([0.6 * np.cos(theta), 0.6 * np.sin(theta), np.zeros_like(theta)])
    u = np.zeros_like(P)
    mats = [material("Film", True, a * 1e-9)]
    mat = np.zeros((nz, ny, nx), dtype=np.int32)
    return {"P": P, "u": u, "mat": mat, "materials": mats, "spacing": a * 1e-9,
            "atom": atom_file(q00_of(hkl, a), a=a)}


def write_run(path, run, h5open, geometry="Bulk", layers=None, phase=None, region=None, limits=None,
              displacement=True, output="diffraction"):
    """Write a FerroDyn run directory: system_setting.json, materials/, data/*.h5, atom file, input.json."""
    os.makedirs(os.path.join(path, "materials"), exist_ok=True)
    os.makedirs(os.path.join(path, "data"), exist_ok=True)
    P, u, mat = run["P"], run["u"], run["mat"]
    nz, ny, nx = P.shape[1:]
    names = [m["Name"] for m in run["materials"]]
    if layers is None:
        layers = ([int(mat.flat[0])], [nz])
    h = run["spacing"]
    setting = {
        "Mesh": {"NX": nx, "NY": ny, "NZ": nz, "DX": h, "DY": h, "DZ": h},
        "Rotation": {"Enabled": False, "EulerAngles": [0.0, 0.0, 0.0]},
        "BoundaryConditions": {"GeometryType": geometry},
        "SimulationControls": {"Temperature": {"Initial": 300.0}},
        "Materials": {"NumberOfMaterials": len(names), "MaterialNames": names,
                      "NumberOfLayers": len(layers[0]), "LayerTypes": list(layers[0]),
                      "LayerSizes": list(layers[1])},
        "Solvers": {},
    }
    with open(os.path.join(path, "system_setting.json"), "w") as fh:
        json.dump(setting, fh)
    for m in run["materials"]:
        with open(os.path.join(path, "materials", m["Name"] + ".json"), "w") as fh:
            json.dump(m, fh)
    with h5open(os.path.join(path, "data", "materialField_0.h5"), "w") as fh:
        fh["materialField"] = np.asarray(mat, dtype=np.int32)
    with h5open(os.path.join(path, "data", "P_Lattice_0.h5"), "w") as fh:
        fh.attrs["Timestep"] = 0
        for c, comp in zip("xyz", P):
            fh["P_Lattice_" + c] = np.asarray(comp, dtype=np.float64)
    cfg = {"platt_file": "data/P_Lattice_0.h5", "output_file": output}
    if displacement:
        with h5open(os.path.join(path, "data", "Displacement_0.h5"), "w") as fh:
            for c, comp in zip("xyz", u):
                fh["Displacement_" + c] = np.asarray(comp, dtype=np.float64)
        cfg["displacement_file"] = "data/Displacement_0.h5"
    if phase is not None:
        with h5open(os.path.join(path, "data", "phase.h5"), "w") as fh:
            for p, comp in enumerate(phase):
                fh["phase_%d" % (p + 1)] = np.asarray(comp, dtype=np.float32)
        cfg["phase_file"] = "data/phase.h5"
    if region is not None:
        with h5open(os.path.join(path, "data", "region.h5"), "w") as fh:
            fh.create_dataset("region", data=np.asarray(region, dtype=np.float32), compression="gzip", shuffle=True)
        cfg["region_file"] = "data/region.h5"
    if limits:
        cfg.update(limits)
    with open(os.path.join(path, "parameter.atom.in"), "w") as fh:
        fh.write(run["atom"])
    with open(os.path.join(path, "input.json"), "w") as fh:
        json.dump(cfg, fh)


def checksum(run):
    """Order-insensitive summary of a run's arrays (sum and sum of squares of each), for fixtures."""
    out = []
    for key in ("P", "u"):
        a = np.asarray(run[key], dtype=np.float64)
        out += [float(a.sum()), float((a * a).sum())]
    out.append(float(np.asarray(run["mat"]).sum()))
    return out
