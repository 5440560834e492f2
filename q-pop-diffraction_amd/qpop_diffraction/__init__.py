"""Q-POP Diffraction for MI355X: kinematic X-ray diffraction from Q-POP-FerroDyn
polarization and displacement fields, computed by hand-written HIP kernels.

The public names mirror the reference package (src/python/__init__.py:6-40).
``write_mgpu_params`` writes the reference's prepared-run file; ``params_run`` runs
such a file on one GPU instead of the reference's MPI program.
"""

from .constants import L0, OS0, PI
from .size_context import SizeContext
from .atom_list import AtomList
from .diffraction import diffraction, reciprocal_centre, q_axes, wavevectors, Engine, Volume
from .io_utils import (
    load_platt_hdf5,
    load_displacement_hdf5,
    load_material_field,
    load_phase_hdf5,
    load_region_hdf5,
    crop_slices,
    write_diffraction_output,
    write_diffraction_axes,
    write_mgpu_params,
    load_diffraction_hdf5,
)

__all__ = [
    "L0", "OS0", "PI",
    "SizeContext",
    "AtomList",
    "diffraction",
    "reciprocal_centre",
    "q_axes",
    "wavevectors",
    "Engine",
    "Volume",
    "load_platt_hdf5",
    "load_displacement_hdf5",
    "load_material_field",
    "load_phase_hdf5",
    "load_region_hdf5",
    "crop_slices",
    "write_diffraction_output",
    "write_diffraction_axes",
    "write_mgpu_params",
    "load_diffraction_hdf5",
]
