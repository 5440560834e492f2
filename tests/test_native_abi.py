"""libqpd.so loads and exports every entry point include/qpd.h declares (CPU, no compute)."""

import ctypes
import os
import re
from pathlib import Path

import numpy as np
import pytest

from qpop_diffraction import _native

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "qpd.h"


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:int|size_t|double|const char\*)\s+(qpd_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_abi():
    names = declared()
    assert set(_native.EXPORTS) == set(names), (names, _native.EXPORTS)


@pytest.mark.skipif(not _native.LIB_PATH.exists(), reason="libqpd.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    for name in declared():
        assert hasattr(lib, name), name


@pytest.mark.skipif(not _native.LIB_PATH.exists(), reason="libqpd.so not built")
def test_host_only_entry_points():
    lib = _native.load_library()
    assert lib.qpd_version() >= 100
    for n, ok in [(1, 1), (2, 1), (104, 1), (4000, 1), (68, 1), (19, 1), (23, 1), (38, 1), (2 * 23, 1), (97, 1),
                  (4001, 1), (65521, 1), (65536, 1), (65537, 0), (65536 * 2, 0)]:
        assert lib.qpd_fft_length_supported(n) == ok, n
    assert lib.qpd_last_error() is not None


@pytest.mark.skipif(not _native.LIB_PATH.exists(), reason="libqpd.so not built")
def test_plan_rejects_bad_arguments():
    """Argument checks run on the host before any device call."""
    lib = _native.load_library()
    handle = ctypes.c_void_p()
    grid = _native.Grid((ctypes.c_int64 * 3)(4, 4, 4), (ctypes.c_int * 3)(0, 0, 1), _native.c_double3(1, 1, 1))
    cr = _native.Crystal()
    cr.n_atom, cr.n_phase, cr.n_struc = 1, 1, 3
    rc = lib.qpd_plan_create(ctypes.byref(handle), ctypes.byref(grid), ctypes.byref(cr), 32)
    assert rc == -1 and b"permutation" in lib.qpd_last_error()
    grid.axis[1] = 1
    grid.axis[2] = 2
    rc = lib.qpd_plan_create(ctypes.byref(handle), ctypes.byref(grid), ctypes.byref(cr), 16)
    assert rc == -1 and b"precision" in lib.qpd_last_error()
    rc = lib.qpd_plan_create(ctypes.byref(handle), ctypes.byref(grid), ctypes.byref(cr), 32)
    assert rc == -1 and b"crystal" in lib.qpd_last_error()  # null tables
    cr.n_struc = 2
    rc = lib.qpd_plan_create(ctypes.byref(handle), ctypes.byref(grid), ctypes.byref(cr), 32)
    assert rc == -2


@pytest.mark.skipif(not _native.LIB_PATH.exists(), reason="libqpd.so not built")
def test_ingest_rejects_bad_arguments():
    """qpd_ingest_fields checks its description on the host before any device call."""
    lib = _native.load_library()
    fake = ctypes.c_void_p(0x1000)
    arg = _native.Ingest((ctypes.c_int64 * 3)(4, 4, 4), 6, 1, (ctypes.c_void_p * 3)(0x1000, 0x1000, 0x1000), 32, 1.0,
                         (ctypes.c_void_p * 3)(None, None, None), 64, 1e-9, 0x1000, 2, 0x1000, 0x1000, 0.4)
    assert lib.qpd_ingest_fields(None, 32, fake, fake, fake, None) == -1
    assert lib.qpd_ingest_fields(ctypes.byref(arg), 16, fake, fake, fake, None) == -1
    arg.z0 = 3  # crop z 3..6 outside a stack of 6
    assert lib.qpd_ingest_fields(ctypes.byref(arg), 32, fake, fake, fake, None) == -1
    assert b"stack" in lib.qpd_last_error()
    arg.z0 = 1
    arg.P_bits = 16
    assert lib.qpd_ingest_fields(ctypes.byref(arg), 32, fake, fake, fake, None) == -1
    arg.P_bits = 32
    arg.U = (ctypes.c_void_p * 3)(0x1000, None, None)
    assert lib.qpd_ingest_fields(ctypes.byref(arg), 32, fake, fake, fake, None) == -1
    assert b"three components" in lib.qpd_last_error()


def test_missing_library_fails_loudly(tmp_path):
    """Without libqpd.so the package raises (with the build command); there is no CPU fallback."""
    import subprocess
    import sys
    code = (
        "import sys\n"
        f"sys.path.insert(0, {str(ROOT / 'q-pop-diffraction_amd')!r})\n"
        "from qpop_diffraction._native import load_library\n"
        "try:\n"
        "    load_library()\n"
        "except ImportError as e:\n"
        "    print('raised:', e)\n"
        "    sys.exit(0)\n"
        "sys.exit(1)\n"
    )
    env = dict(os.environ, QPD_LIB=str(tmp_path / "absent" / "libqpd.so"))
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "build" in r.stdout


def test_no_gpu_fails_loudly(tmp_path):
    """On a machine without a GPU the drop-in diffraction() raises instead of computing elsewhere."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import golden_cases
    from helpers import oracle_inputs
    from qpop_diffraction import diffraction
    from qpop_diffraction.h5lite import open_file
    golden_cases.case_bulk(open_file, str(tmp_path))
    params, atoms, oPhase, oStruc, u, region, _ = oracle_inputs(tmp_path)
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        diffraction(params, atoms, oPhase, oStruc, u, region)
