"""Drop-in for the reference's backend/pm_kernel.py.

Same public names and arguments: build_force_kernels, grid_fuer,
pm_force_field, gather_accel, pm_accelerations, _cic_deposit, _cic_gather,
NBodyPM (pm_kernel.py:34-293).  Device arrays are torch tensors on the ROCm
device (the reference passes CuPy arrays); the FFTs, deposits and gathers
run in libssim.
"""
try:
    from . import _path  # noqa: F401
except ImportError:  # imported as a top-level module from this directory
    import _path  # noqa: F401

from solarsystemsimulator_amd.constants import G_AU  # noqa: F401
from solarsystemsimulator_amd.pm import (  # noqa: F401
    NBodyPM, _cic_deposit, _cic_gather, build_force_kernels, gather_accel,
    grid_fuer, pm_accelerations, pm_force_field)
