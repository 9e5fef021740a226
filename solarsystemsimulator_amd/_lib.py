"""ctypes binding of libssim.so (the C-ABI declared in include/ssim.h).

The library is loaded from the package directory, where
``python -m solarsystemsimulator_amd.build`` puts it.  There is no fallback:
if the shared object is missing or does not load, every GPU entry point
raises, so a run can never silently compute on the host.

Device buffers are torch tensors on the ROCm device; only their
``data_ptr()`` crosses the boundary, together with the current HIP stream.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import c_double, c_float, c_int, c_longlong, c_size_t, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libssim.so"

SSIM_OK = 0
SSIM_EINVAL = -1
SSIM_EHIP = -2

# largest n of ssim_energy_noalloc (SSIM_ENERGY_NOALLOC_MAX_N in ssim.h)
SSIM_ENERGY_NOALLOC_MAX_N = 1 << 24

PREC_F64 = 0
PREC_F32C = 1
SCHEME_KDK = 0
SCHEME_YOSHIDA4 = 1

_P = c_void_p  # device pointer / stream

# name -> (restype, argtypes).  Kept in the order of include/ssim.h.
SIGNATURES = {
    "ssim_last_error": (ctypes.c_char_p, []),
    "ssim_version": (ctypes.c_char_p, []),
    "ssim_device_count": (c_int, []),
    "ssim_timing": (c_int, [c_int]),
    "ssim_timing_report": (c_int, [ctypes.c_char_p, _P, _P]),
    "ssim_allpairs_work_bytes": (c_size_t, [c_int, c_int, c_int, c_int]),
    "ssim_allpairs": (c_int, [c_int, c_int, c_int, _P, _P, _P, _P, c_int, _P,
                              _P, _P, _P, _P, _P, c_double, c_double, _P,
                              c_size_t, _P]),
    "ssim_selfgrav_work_bytes": (c_size_t, [c_int, c_int, c_int, c_int]),
    "ssim_selfgrav_run": (c_int, [c_int, c_int, c_int, c_int, c_int,
                                  _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                  c_double, c_double, c_double, c_int, c_int,
                                  _P, c_longlong, _P, c_size_t, _P]),
    "ssim_kick": (c_int, [c_int, c_int, _P, _P, _P, _P, _P, _P, c_double, _P]),
    "ssim_drift": (c_int, [c_int, c_int, _P, _P, _P, _P, _P, _P, c_double, _P]),
    "ssim_drift_periodic": (c_int, [c_int, c_int, _P, _P, _P, _P, _P, _P, c_double,
                                    _P, c_double, _P]),
    "ssim_kick_drift": (c_int, [c_int, c_int] + [_P] * 9 + [c_double, c_double, _P]),
    "ssim_kick_snap": (c_int, [c_int, c_int, c_int, _P, _P, _P, _P, _P, _P, _P,
                               _P, _P, c_double, _P, _P]),
    "ssim_fill_f32": (c_int, [c_longlong, _P, c_float, _P]),
    "ssim_scatter_f64": (c_int, [c_int, _P, _P, _P, c_int, _P]),
    "ssim_ring_slot_bytes": (ctypes.c_longlong, [c_int, c_int, c_int]),
    "ssim_ring_work_bytes": (c_size_t, [c_int, c_int]),
    "ssim_frame_rows": (c_int, [c_int, c_int, c_int, c_int, _P, _P, _P, _P, _P, _P]),
    "ssim_ring_slots": (c_int, [c_int] * 6 + [_P] * 8 + [_P, c_size_t, _P]),
    "ssim_nbody_work_bytes": (c_size_t, [c_int]),
    "ssim_nbody_frame": (c_int, [c_int] + [_P] * 9 + [c_int] + [_P] * 9 +
                         [c_double, c_double, c_double, c_int, c_double, c_int,
                          c_double, c_double, c_int] + [_P] * 6 +
                         [_P, c_size_t, _P]),
    "ssim_nbody_frame3d": (c_int, [c_int] + [_P] * 12 + [c_int] + [_P] * 12 +
                           [c_double, c_double, c_double, c_int, c_double, c_int,
                            c_double, c_double, c_int] + [_P] * 6 +
                           [_P, c_size_t, _P]),
    "ssim_energy_work_bytes": (c_size_t, [c_int, c_int]),
    "ssim_energy": (c_int, [c_int, c_int, _P, _P, _P, _P, _P, _P, _P, c_double,
                            c_double, _P, _P, c_size_t, _P]),
    "ssim_energy_noalloc": (c_int, [c_int, c_int, _P, _P, _P, _P, _P, _P, _P,
                                    c_double, c_double, _P, _P]),
    "ssim_fft_work_bytes": (c_size_t, [c_int, _P, c_int]),
    "ssim_fft_prepare": (c_int, [c_int, _P, c_int, _P, c_size_t, _P]),
    "ssim_fft_r2c": (c_int, [c_int, _P, c_int, _P, _P, _P, c_size_t, _P]),
    "ssim_fft_c2r": (c_int, [c_int, _P, c_int, _P, _P, _P, c_size_t, _P]),
    "ssim_cic_deposit": (c_int, [c_int, c_int, c_int, _P, _P, _P, _P, _P, _P,
                                 c_double, c_int, _P, _P, _P]),
    "ssim_cic_deposit_tiles_work_bytes": (c_size_t, [c_int, c_int, _P]),
    "ssim_cic_deposit_tiles": (c_int, [c_int, c_int, c_int, _P, _P, _P, _P, _P, _P,
                                       c_double, c_int, _P, _P, _P, c_size_t, _P]),
    "ssim_cic_gather": (c_int, [c_int, c_int, c_int, _P, _P, _P, _P, _P,
                                c_double, c_int, c_int, _P, _P, _P, _P, _P, _P,
                                _P]),
    "ssim_minmax_work_bytes": (c_size_t, [c_int]),
    "ssim_minmax": (c_int, [c_int, c_int, _P, _P, _P, _P, _P, c_size_t, _P]),
    "ssim_box_after_kdk": (c_int, [c_int, c_int] + [_P] * 9 + [c_double, c_double, _P,
                                                             _P, c_size_t, _P]),
    "ssim_kick_box_work_bytes": (c_size_t, [c_int]),
    "ssim_kick_box": (c_int, [c_int, c_int] + [_P] * 9 + [c_double, c_double, _P, _P,
                                                        c_size_t, _P]),
    "ssim_kicks_drift": (c_int, [c_int, c_int] + [_P] * 9 + [c_double, c_double, _P]),
    "ssim_kick_drift_box": (c_int, [c_int, c_int] + [_P] * 9 + [c_double, c_double, _P,
                                                              _P, c_size_t, _P]),
    "ssim_pm_force_kernel": (c_int, [c_int, c_int, c_int, c_double, c_double,
                                     c_double, c_double, c_double, c_int, _P,
                                     _P]),
    "ssim_pm_deconvolve": (c_int, [c_int, c_int, c_int, c_int, _P, _P]),
    "ssim_tsc_deposit": (c_int, [c_int, _P, _P, _P, _P, c_int, _P, c_double,
                                 c_int, _P, _P]),
    "ssim_tsc_deposit_work_bytes": (c_size_t, [c_int]),
    "ssim_tsc_deposit_binned": (c_int, [c_int, _P, _P, _P, _P, c_int, _P, c_double,
                                        _P, _P, c_int, _P, _P, c_size_t, _P]),
    "ssim_tsc_deposit_blocked": (c_int, [c_int, _P, c_double, c_int, _P, _P, c_int,
                                         _P, _P]),
    "ssim_tsc_deposit_tiles": (c_int, [c_int, _P, c_double, c_int, _P, _P, c_double,
                                       c_int, _P, _P]),
    "ssim_tsc_gather": (c_int, [c_int, _P, _P, _P, _P, c_int, _P, c_double, c_int,
                                _P, _P, _P, _P, _P, _P, c_int, _P]),
    "ssim_tsc_gather_corrected": (c_int, [c_int, _P, _P, _P, _P, c_int, _P, c_double,
                                          _P, _P, _P, _P, c_int, _P, c_float, c_int,
                                          _P, _P, _P, c_int, _P]),
    "ssim_tsc_gather_box_work_bytes": (c_size_t, [c_int]),
    "ssim_tsc_gather_box": (c_int, [c_int, _P, _P, _P, _P, c_int, _P, c_double, _P, _P, _P,
                                    _P, c_int, _P, c_float, c_int, _P, _P, _P, c_int,
                                    c_double, c_double, _P, _P, c_size_t, _P]),
    "ssim_spectrum_mul": (c_int, [c_longlong, c_int, _P, _P, _P, _P]),
    "ssim_spectrum_imag": (c_int, [c_longlong, c_int, _P, _P, _P]),
    "ssim_pm_conv_isolated3": (c_int, [c_int, c_int] + [_P] * 9 + [_P, c_size_t,
                                                                  _P]),
    "ssim_pm_potential_periodic": (c_int, [_P, c_int, _P, _P, _P, c_double, c_double, _P,
                                           c_size_t, _P]),
    "ssim_pm_poisson": (c_int, [c_int, _P, c_double, c_double, c_int, _P, _P]),
    "ssim_pm_gradient": (c_int, [c_int, _P, c_double, c_int, _P, _P, _P, _P,
                                 _P]),
    "ssim_pm_pad": (c_int, [c_int, c_int, c_int, _P, _P, c_int, _P]),
    "ssim_cells_work_bytes": (c_size_t, [c_int, c_int]),
    "ssim_cells_build": (c_int, [c_int, _P, _P, _P, _P, _P, c_double, _P, _P,
                                 _P, _P, _P, c_size_t, _P]),
    "ssim_cells_bin_work_bytes": (c_size_t, [c_int, c_int]),
    "ssim_cells_bin": (c_int, [c_int, _P, _P, _P, _P, c_double, _P, _P, _P, _P,
                               c_size_t, _P]),
    "ssim_split_bin_work_bytes": (c_size_t, [c_int, c_int, c_int]),
    "ssim_split_bin": (c_int, [c_int, _P, _P, _P, _P, c_int, _P, c_double, c_int, _P,
                               _P, _P, _P, _P, _P, _P, _P, c_size_t, _P]),
    "ssim_split_bin_all_work_bytes": (c_size_t, [c_int, c_int, c_int]),
    "ssim_split_bin_all": (c_int, [c_int, c_int, _P, _P, _P, _P, c_int, _P, c_double, c_int,
                                   _P, _P, _P, _P, _P, _P, _P, _P, _P, c_size_t, _P]),
    "ssim_pp_work_bytes": (c_size_t, [c_int, c_int]),
    "ssim_pp_short": (c_int, [c_int, _P, _P, _P, _P, c_double, _P, _P, _P, _P,
                              _P, c_double, c_double, c_double, c_double, _P,
                              _P, _P, c_int, _P, c_size_t, _P]),
    "ssim_pp_near": (c_int, [c_int, _P, _P, _P, _P, c_double, _P, _P, _P, _P, _P,
                             c_double, c_double, c_double, c_double, _P, _P, _P, c_int,
                             _P, c_size_t, _P]),
    "ssim_pp_near_hilo": (c_int, [c_int, _P, _P, _P, _P, c_double, _P, _P, _P, _P, _P,
                                  _P, c_double, c_double, c_double, c_double, _P, _P, _P,
                                  c_int, _P, c_size_t, _P]),
    "ssim_pp_near_bits": (c_int, [c_int, _P, _P, _P, _P, c_double, _P, _P, _P, _P, _P,
                                  _P, c_double, c_double, c_double, c_double, _P, _P, _P,
                                  c_int, c_int, _P, c_size_t, _P]),
    "ssim_pp_count": (c_int, [c_int, _P, _P, _P, _P, c_double, _P, _P, _P, _P,
                              _P, c_double, c_double, c_double, _P, _P, c_size_t,
                              _P]),
    "ssim_speed_percentile_work_bytes": (c_size_t, [c_int]),
    "ssim_speed_percentile": (c_int, [c_int, _P, _P, _P, _P, c_float, _P, _P, c_size_t,
                                      _P]),
    "ssim_collide_batch": (c_int, [c_longlong]),
    "ssim_collide_work_bytes": (c_size_t, [c_int, c_longlong]),
    "ssim_collide_hits": (c_int, [c_int, _P, _P, _P, _P, _P, _P, _P, c_float,
                                  c_float, c_float, c_double, c_double,
                                  c_longlong, c_longlong, _P, _P, _P, _P,
                                  c_size_t, _P]),
    "ssim_merge_work_bytes": (c_size_t, [c_int, c_int]),
    "ssim_merge_scan": (c_int, [c_int] + [_P] * 9 + [c_int, _P, c_float, c_int,
                                c_float, c_float, c_float, c_longlong, _P, _P,
                                _P, c_longlong, _P, _P, _P, c_size_t, _P]),
    "ssim_bounce_work_bytes": (c_size_t, [c_int, c_int]),
    "ssim_bounce_deltas": (c_int, [c_int, c_int] + [_P] * 8 + [c_double, c_double]
                           + [_P] * 7 + [_P, c_size_t, _P]),
    "ssim_add_body_deltas": (c_int, [c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "ssim_apply_merges": (c_int, [c_int, _P, c_int, _P, _P, _P, c_int, _P, _P, _P, _P,
                                  _P, _P, _P, c_int, c_int, c_double, c_double,
                                  c_double, _P, _P, _P, _P, _P, _P, _P, _P]),
    "ssim_mlp_param_count": (c_int, []),
    "ssim_mlp_to_bf16": (c_int, [_P, _P, _P]),
    "ssim_mlp_features": (c_int, [c_int, c_int, _P, _P, _P, _P, c_int, _P,
                                  c_double, c_int, _P, c_double, _P, _P]),
    "ssim_mlp_forward": (c_int, [c_int, _P, _P, _P, c_float, c_int, _P, _P, _P,
                                 _P, _P, _P, _P, c_int, _P, _P, _P]),
    "ssim_mlp_residual": (c_int, [c_int, _P, _P, _P, _P, _P, _P, c_double, c_int,
                                  c_double, _P, _P]),
    "ssim_mlp_mse_grad": (c_int, [c_int, _P, _P, _P, _P, _P]),
    "ssim_mlp_train_work_bytes": (c_size_t, [c_int]),
    "ssim_mlp_backward": (c_int, [c_int, _P, _P, _P, _P, _P, _P, _P, c_size_t,
                                  _P]),
    "ssim_mlp_adam": (c_int, [_P, _P, _P, _P, c_float, c_float, c_float,
                              c_float, c_int, _P, _P]),
}


class SsimError(RuntimeError):
    """A HIP runtime failure reported by libssim."""


_lib = None


def load() -> ctypes.CDLL:
    """Open libssim.so once and attach the signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with "
            "`python -m solarsystemsimulator_amd.build`")
    lib = ctypes.CDLL(os.fspath(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    """Turn a libssim status code into ValueError / SsimError."""
    if rc == SSIM_OK:
        return
    msg = load().ssim_last_error().decode(errors="replace")
    if rc == SSIM_EINVAL:
        raise ValueError(msg)
    raise SsimError(msg or f"libssim error {rc}")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))


def int_array(vals):
    """A host int32 array for `const int*` arguments (kept alive by caller)."""
    arr = (ctypes.c_int * len(vals))(*[int(v) for v in vals])
    return arr


def double_array(vals):
    return (ctypes.c_double * len(vals))(*[float(v) for v in vals])


def ptr(t) -> int | None:
    """Device address of a torch tensor (None for None)."""
    return None if t is None else t.data_ptr()


def stream_handle(device=None) -> int:
    """The current torch stream of `device` as a hipStream_t value."""
    import torch
    return torch.cuda.current_stream(device).cuda_stream
