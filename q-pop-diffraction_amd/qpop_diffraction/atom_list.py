"""The crystal of ``parameter.atom.in``: lattice, atoms, form factors, P -> shift coefficients.

Mirrors ``AtomList`` of the reference (src/python/atom_list.py:9-193).  The
file is a stream of numbers: text after ``!`` is a comment, line breaks carry
no meaning and tokens that are not numbers are skipped (atom_list.py:72-76).
Order: q00 (3), aC (3x3, one lattice vector per row, nm), nAtom, nAtom x 3
reduced coordinates, then per phase: its 1-based id, nAtom (Re f, Im f), and
per atom nStruc rows of (su1, su2, su3) in m per C/m^2.
"""

import numpy as np

from .constants import L0, OS0


def _numbers(text):
    """All numeric tokens of the file, comments removed."""
    values = []
    for line in text.splitlines():
        body = line.split("!", 1)[0]
        for token in body.split():
            try:
                values.append(float(token))
            except ValueError:
                pass
    return values


class AtomList:
    """Unit cell of the diffraction model."""

    def __init__(self):
        self.nPhase = 0
        self.nStruc = 0
        self.q00 = np.zeros(3)
        self.aC = np.zeros((3, 3))
        self.xAtom = None
        self.fAtom = None
        self.fAtomR = None
        self.fAtomI = None
        self.su1 = self.su2 = self.su3 = None
        self.nAtom = 0

    def read(self, filename, nPhase, nStruc):
        """Parse the file; False when it cannot be opened or nAtom <= 0 (atom_list.py:37-136)."""
        self.nPhase, self.nStruc = nPhase, nStruc
        try:
            with open(filename, "r") as fh:
                text = fh.read()
        except IOError:
            print(f"Cannot open file: {filename}")
            return False

        values = _numbers(text)
        cursor = [0]

        def take():
            # Running out of numbers is an IndexError, as indexing the token list is.
            value = values[cursor[0]]
            cursor[0] += 1
            return value

        self.q00 = np.array([take() for _ in range(3)])
        self.aC = np.array([[take() for _ in range(3)] for _ in range(3)])
        self.nAtom = int(take())
        if self.nAtom <= 0:
            print(f"Invalid number of atoms: {self.nAtom}")
            return False

        print("Reading atom coordinates")
        self.xAtom = np.array([[take() for _ in range(3)] for _ in range(self.nAtom)])

        print("Reading atom form factors")
        self.fAtom = np.zeros((nPhase, self.nAtom), dtype=complex)
        self.fAtomR = np.zeros((nPhase, self.nAtom))
        self.fAtomI = np.zeros((nPhase, self.nAtom))
        su = np.zeros((3, nStruc, nPhase, self.nAtom))
        for _ in range(nPhase):
            p = int(take()) - 1
            print(f"Read input parameters for phase # {p + 1}")
            for a in range(self.nAtom):
                self.fAtomR[p, a] = take()
                self.fAtomI[p, a] = take()
                self.fAtom[p, a] = complex(self.fAtomR[p, a], self.fAtomI[p, a])
            for a in range(self.nAtom):
                for k in range(nStruc):
                    for d in range(3):
                        su[d, k, p, a] = take()
        self.su1, self.su2, self.su3 = su[0], su[1], su[2]
        return True

    def validate(self):
        """Print what was read (atom_list.py:138-174)."""
        print("Validating read atom parameters")
        print(f"q00: {self.q00}")
        print("aC:")
        for row in self.aC:
            print(f"  {row}")
        print(f"nAtom: {self.nAtom}")
        print("xAtom:")
        for row in self.xAtom:
            print(f"  {row}")
        print("fAtom (per phase):")
        for p in range(self.nPhase):
            print("  " + " ".join(f"({self.fAtomR[p, a]},{self.fAtomI[p, a]})" for a in range(self.nAtom)))
        return True

    def setup(self):
        """Convert su from m per C/m^2 to nm per C/m^2 (atom_list.py:176-193)."""
        print("Setting up atom parameters")
        scale = L0 / OS0
        self.su1 = self.su1 / scale
        self.su2 = self.su2 / scale
        self.su3 = self.su3 / scale
        return True
