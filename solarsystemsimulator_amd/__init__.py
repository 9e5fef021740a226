"""MI355X-native GPU physics backend of the solar-system / self-gravity
simulator: hand-written gfx950 HIP kernels in libssim.so behind a ctypes
C-ABI, with drop-in modules carrying the reference's module, class and
function names under ``solarsystemsimulator_amd/backend``.
"""
from .constants import G_AU  # noqa: F401

__version__ = "0.1.0"
