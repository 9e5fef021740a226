"""Drop-in for the reference's backend/nbody_kernel.py.

Same public names: NBodyCuda, pick_device, pick_devices,
pick_detect_devices, f64_score and the module constants M_MAX, K_MAX,
G_AU, SOFTENING, MAX_SUB_DT_YEARS, MAX_SUB_STEPS_PER_FRAME, YOSHIDA_W1,
YOSHIDA_W0, SUB_SAMPLES, SUB_SAMPLES_MAX (nbody_kernel.py:58-72,563-1060).
The raster integrator runs in libssim (HIP, gfx950) on one GPU.
"""
try:
    from . import _path  # noqa: F401
except ImportError:  # imported as a top-level module from this directory
    import _path  # noqa: F401

from solarsystemsimulator_amd.constants import (  # noqa: F401
    G_AU, G_MAX, K_MAX, M_MAX, MAX_SUB_DT_YEARS, MAX_SUB_STEPS_PER_FRAME,
    SOFTENING, SUB_SAMPLES, SUB_SAMPLES_MAX, YOSHIDA_W0, YOSHIDA_W1)
from solarsystemsimulator_amd.nbody import (  # noqa: F401
    NBodyCuda, pick_detect_devices, pick_device, pick_devices)


def f64_score(i: int) -> float:
    """Relative FP64 throughput of device i (nbody_kernel.py:569-572): CU
    count x clock; every MI355X counts the same."""
    import torch
    p = torch.cuda.get_device_properties(i)
    return float(p.multi_processor_count)
