"""Unit scales shared by the host code (reference: src/python/constants.py:8-12).

Lengths are carried in nanometres inside the package, while FerroDyn files hold
metres; the order parameter (lattice polarization) is in C/m^2 with unit scale.
"""

import math

L0 = 1e-9      # metres per nanometre
OS0 = 1.0      # order-parameter unit (C/m^2)
PI = math.pi
