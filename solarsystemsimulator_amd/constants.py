"""Module constants of the reference physics backend, same names and values.

nbody_kernel.py:58-72, selfgrav_kernel.py:75-78, gpu_verbund.py:30-31.
"""
import numpy as np

M_MAX = 64                      # massive bodies per nbody state
K_MAX = 16                      # raster samples per step_batch call
G_AU = 4 * np.pi * np.pi        # AU^3 / (M_sun yr^2)
SOFTENING = 1e-6                # nbody: added to r^2 (acts as eps^2)
MAX_SUB_DT_YEARS = 0.5 / 365.25
MAX_SUB_STEPS_PER_FRAME = 100_000
YOSHIDA_W1 = 1.0 / (2.0 - 2.0 ** (1.0 / 3.0))
YOSHIDA_W0 = -(2.0 ** (1.0 / 3.0)) / (2.0 - 2.0 ** (1.0 / 3.0))
SUB_SAMPLES = 8
SUB_SAMPLES_MAX = 32

G_MAX = 8                       # device-list limit kept for API checks
BLOCK = 256                     # selfgrav tile width (kept for callers)
KALIBRIER_N = 4096
KALIBRIER_SCHRITTE = 12
KALIBRIER_RUNDEN = 3

# film producer (film_producer.py:36-78)
EV_BYTES = 32                   # event record: "<dIIfIff"
EV_FORMAT = "<dIIfIff"          # t_days f64 | a u32 | b u32 | mass f32 | kind u32 | x f32 | y f32
BOUNCE_E = 0.6                  # restitution of asteroid bounces
SUB_MAX_ANTEIL = 32             # one sub-sample track slot per this many bodies
SUB_MAX_MIN = 2048
SUB_SAMPLES_DEFAULT = 8
M_SUB_MAX = 32
