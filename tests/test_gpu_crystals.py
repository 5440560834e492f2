"""Cells beyond the five-atom perovskite: more atoms, source groups, classes and phases.

The golden cases all use five-atom PbTiO3-like cells (two source groups, two z-classes).
These runs exercise the rest of the grouping logic against the oracle, through the CLI:
  * groups9: nine atoms in three groups of equal P -> shift tensors, on two z planes: pass 0
    transforms three group representatives and pass 1 mixes them into two z-classes;
  * classes7: seven atoms with distinct shift tensors on four z planes: the class form;
  * phases3: three ferroelectric phases (one-hot from the material field) whose form factors
    and shift tensors differ per phase, with form-factor ratios shared by a subset of atoms;
  * fractions3: the same crystal with fractional phase weights from a phase file.
Full 3 x 3 shift tensors, complex form factors and fractional positions on a quarter grid.
FP64 is held to the reference's parity bound (rtol 1e-9, 1e-12 of the peak), FP32 to the bulk
bound (1e-5 of the peak, the peak on the same voxel).
"""

import numpy as np
import pytest

import golden_cases
from helpers import check_fp32, oracle_run, run_cli
from qpop_diffraction.h5lite import open_file

pytestmark = pytest.mark.gpu

synth = golden_cases.synth
A = 0.4  # nm
NX, NY, NZ = 24, 20, 16
HKL = (1, 0, 2)


def crystal_text(pos, forms, su):
    """parameter.atom.in: q00, aC, the atoms' fractional positions, then per phase its number,
    every atom's complex form factor and every atom's three rows of P -> shift (m per C/m^2)."""
    num = lambda v: repr(float(v))
    q00 = [2 * np.pi * h / A for h in HKL]
    lines = ["! test cell", " ".join(num(v) for v in q00)]
    lines += [" ".join(num(v) for v in row) for row in np.diag([A, A, A])]
    lines += [str(len(pos))] + [" ".join(num(v) for v in p) for p in pos]
    for ph in range(len(forms)):
        lines.append(str(ph + 1))
        lines += [f"{num(f.real)} {num(f.imag)}" for f in forms[ph]]
        for a in range(len(pos)):
            lines += [" ".join(num(v) for v in row) for row in su[ph][a]]
    return "\n".join(lines) + "\n"


def cell(name, rng):
    """(positions, forms[phase][atom], su[phase][atom] 3x3, number of phases)."""
    quarter = lambda n, zs: [(rng.integers(0, 4) / 4, rng.integers(0, 4) / 4, zs[rng.integers(0, len(zs))])
                             for _ in range(n)]
    cform = lambda n: rng.uniform(5, 60, n) + 1j * rng.uniform(0, 3, n)
    tensor = lambda: rng.normal(0, 2e-11, (3, 3))
    if name == "groups9":
        pos = quarter(9, (0.0, 0.5))
        group = [0, 0, 0, 1, 1, 1, 2, 2, 2]
        t = [tensor() for _ in range(3)]
        return pos, [cform(9)], [[t[g] for g in group]], 1
    if name == "classes7":
        pos = quarter(7, (0.0, 0.25, 0.5, 0.75))
        return pos, [cform(7)], [[tensor() for _ in range(7)]], 1
    # three phases: atoms 0-2 share their shift tensor and their form-factor ratios across phases
    # (three source groups), on four y planes (four y-classes) and two z planes
    pos = [(0.0, 0.0, 0.0), (0.5, 0.25, 0.5), (0.25, 0.5, 0.0), (0.75, 0.75, 0.5), (0.5, 0.0, 0.0)]
    base, t = cform(5), tensor()
    forms, su = [], []
    for ph in range(3):
        f = base * np.concatenate([np.full(3, 1.0 + 0.2 * ph), rng.uniform(0.7, 1.3, 2)])
        forms.append(f)
        su.append([t * (1.0 + 0.1 * ph)] * 3 + [tensor(), tensor()])
    return pos, forms, su, 3


def write_case(path, name, seed=7):
    rng = np.random.default_rng(seed)
    pos, forms, su, nph = cell(name, rng)
    z, y, x = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    P = np.stack([0.3 * np.sin(2 * np.pi * x / 12), 0.2 * np.cos(2 * np.pi * y / 10),
                  0.5 * np.sign(np.sin(2 * np.pi * (x + y) / 16) + 0.1)]) + rng.normal(0, 0.02, (3, NZ, NY, NX))
    u = rng.normal(0, 2e-12, (3, NZ, NY, NX))
    mats = [synth.material(f"Phase{p + 1}", True, A * 1e-9) for p in range(nph)]
    mat = ((x // 8) % nph).astype(np.int32)
    run = golden_cases.plant_run(P, u, mat=mat, materials=mats, atom=crystal_text(pos, forms, su))
    run["spacing"] = A * 1e-9
    phase = None
    if name == "fractions3":
        w = rng.uniform(0.2, 1.0, (3, NZ, NY, NX))
        phase = w / w.sum(axis=0)
    synth.write_run(str(path), run, open_file, phase=phase)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("name", ["groups9", "classes7", "phases3", "fractions3"])
def test_cells_against_oracle(tmp_path, name, precision):
    print(f"cell {name} precision {precision}", flush=True)
    write_case(tmp_path, name)
    I, _ = run_cli(tmp_path, precision=precision)
    ref, _ = oracle_run(tmp_path)
    if precision == 64:
        np.testing.assert_allclose(I, ref, rtol=1e-9, atol=1e-12 * ref.max())
    else:
        check_fp32(I, ref, name)


def workspace_bytes(path, precision=64):
    """Workspace of the plan the CLI makes for the run (file order: x fastest): 3 fields per
    source group in pass 0, or 3 per y-class."""
    from helpers import oracle_inputs
    from qpop_diffraction import Engine
    params, atoms, *_ = oracle_inputs(path)
    return Engine(params, atoms).plan((NX, NY, NZ), (0, 1, 2), precision).workspace_bytes


@pytest.mark.parametrize("name", ["groups9", "phases3"])
def test_groups_match_class_form(tmp_path, name, monkeypatch):
    """Cells with shared shift tensors take the source-group plan (a smaller workspace than the
    class form's), and the two forms agree."""
    write_case(tmp_path, name)
    grouped = workspace_bytes(tmp_path)
    I_groups, _ = run_cli(tmp_path, precision=64)
    monkeypatch.setenv("QPD_CLASS_FORM", "1")
    assert grouped < workspace_bytes(tmp_path), "the cell did not take the source-group plan"
    I_class, _ = run_cli(tmp_path, precision=64)
    np.testing.assert_allclose(I_groups, I_class, rtol=1e-9, atol=1e-12 * I_class.max())
