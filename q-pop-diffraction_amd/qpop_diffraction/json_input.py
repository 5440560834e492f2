"""``input.json`` of a diffraction run (reference: src/python/json_input.py:22-113).

Keys: platt_file (required), displacement_file, phase_file, region_file,
x_limits / y_limits / z_limits ([start, end], 0-based inclusive) and
output_file (default "diffraction").  Missing files raise FileNotFoundError,
malformed values ValueError; other keys are ignored.
"""

from __future__ import annotations

import json
from pathlib import Path


def _existing(data, key, must_be_file):
    if data.get(key) is None:
        return None
    path = Path(data[key])
    ok = path.is_file() if must_be_file else path.exists()
    if not ok:
        raise FileNotFoundError(f"{key} not found: {path}")
    return path


def _limits(key, raw):
    if raw is None:
        return None
    if not isinstance(raw, (list, tuple)) or len(raw) != 2:
        raise ValueError(f"'{key}' must be a 2-element [start, end] array, got: {raw!r}")
    start, end = int(raw[0]), int(raw[1])
    if start < 0:
        raise ValueError(f"'{key}' start index must be >= 0, got {start}.")
    if end < start:
        raise ValueError(f"'{key}' end ({end}) must be >= start ({start}).")
    return [start, end]


def load_input_json(path: str | Path = "input.json") -> dict:
    """Parse and validate input.json; every key is present in the result."""
    path = Path(path)
    if not path.is_file():
        raise FileNotFoundError(f"Input JSON not found: {path}")
    with path.open("r", encoding="utf-8") as fh:
        try:
            data = json.load(fh)
        except json.JSONDecodeError as exc:
            raise ValueError(f"Failed to parse {path}: {exc}") from exc

    if "platt_file" not in data:
        raise ValueError("input.json must contain 'platt_file'.")
    platt = Path(data["platt_file"])
    if not platt.exists():
        raise FileNotFoundError(f"platt_file not found: {platt}")
    displacement = _existing(data, "displacement_file", must_be_file=False)
    if data.get("phaseFra_file") is not None:
        raise ValueError("'phaseFra_file' (text) was replaced by 'phase_file' "
                         "(HDF5 datasets phase_1 .. phase_<nPhase>).")
    return {
        "platt_file": platt,
        "displacement_file": displacement,
        "phase_file": _existing(data, "phase_file", must_be_file=True),
        "region_file": _existing(data, "region_file", must_be_file=True),
        "x_limits": _limits("x_limits", data.get("x_limits")),
        "y_limits": _limits("y_limits", data.get("y_limits")),
        "z_limits": _limits("z_limits", data.get("z_limits")),
        "output_file": str(data.get("output_file", "diffraction")),
    }
