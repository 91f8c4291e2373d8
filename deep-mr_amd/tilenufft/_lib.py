"""ctypes binding of libtilenufft.so (the C ABI declared in include/tilenufft.h).

This replaces deepmr's Numba FFI (src/deepmr/_utils/backend.py:9-61): tensors
cross the boundary as raw device pointers plus sizes, kernels run on torch's
current HIP stream, and errors come back as status codes that are raised here
as ``RuntimeError`` with the library's message. There is no fallback: if the
library is missing the import of the compute path fails loudly.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libtilenufft.so")

c_int32, c_int64, c_float, c_void_p = ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_void_p


class PlanParams(ctypes.Structure):
    """Mirror of ``tn_plan_params`` (per-axis arrays in (z, y, x) order)."""

    _fields_ = [
        ("ndim", c_int32),
        ("nframes", c_int32),
        ("npts", c_int64),
        ("shape", c_int32 * 3),
        ("grid", c_int32 * 3),
        ("width", c_int32 * 3),
        ("beta", c_float * 3),
        ("shift", c_int32 * 3),
        ("tile", c_int32 * 3),
        ("itile", c_int32 * 3),
        ("interp_mode", c_int32),
        ("interp_chunk", c_int32),
        ("grid_chunk", c_int32),
        ("grid_split", c_int32),
        ("grid_mode", c_int32),
        ("inv_scale", c_float),
        ("apod", c_void_p * 3),
    ]


# name -> (restype, argtypes)
_SIGNATURES = {
    "tn_plan_create": (c_int32, [ctypes.POINTER(PlanParams), c_void_p, c_void_p, ctypes.POINTER(c_void_p)]),
    "tn_plan_destroy": (c_int32, [c_void_p]),
    "tn_plan_device_bytes": (c_int64, [c_void_p]),
    "tn_plan_entries": (c_int64, [c_void_p]),
    "tn_plan_footprint_cells": (c_int64, [c_void_p]),
    "tn_plan_tile": (c_int32, [c_void_p, ctypes.POINTER(c_int32)]),
    "tn_plan_grid_mode": (c_int32, [c_void_p]),
    "tn_cg_grad_work_floats": (c_int64, [c_int64, c_int64]),
    "tn_cg_iterate": (c_int32, [c_void_p] * 11 + [c_int64, c_int64, c_void_p]),
    "tn_cg_iterate_backward": (c_int32, [c_void_p] * 15 + [c_int64, c_int64, c_void_p]),
    "tn_coil_combine": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int64, c_int64,
                                  c_void_p]),
    "tn_plan_prepare": (c_int32, [c_void_p, c_int64]),
    "tn_interp": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p]),
    "tn_grid": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p]),
    "tn_grid_static": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p]),
    "tn_interp_lowrank": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_int32, c_int32, c_void_p,
                                    c_void_p, c_void_p]),
    "tn_grid_lowrank": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_int32, c_int32, c_void_p,
                                  c_void_p, c_void_p]),
    "tn_apod_pad": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
    "tn_crop_apod": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
    "tn_plan_has_fft": (c_int32, [c_void_p]),
    "tn_fft_forward": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
    "tn_fft_adjoint": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
    "tn_fft_z_forward": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
    "tn_fft_z_adjoint": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
    "tn_batch_create": (c_int32, [ctypes.POINTER(c_void_p), ctypes.POINTER(c_int64), ctypes.POINTER(c_int64),
                                  ctypes.POINTER(c_int64), ctypes.POINTER(c_int64), c_int32,
                                  ctypes.POINTER(c_void_p)]),
    "tn_batch_destroy": (c_int32, [c_void_p]),
    "tn_batch_interp": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "tn_batch_grid": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "tn_sense_expand": (c_int32, [c_void_p, c_int64, c_int64, c_void_p, c_void_p]),
    "tn_sense_combine": (c_int32, [c_void_p, c_int64, c_int64, c_void_p, c_void_p]),
    "tn_cg_work_floats": (c_int64, [c_int64, c_int64]),
    "tn_cg_step": (c_int32, [c_void_p, c_int64, c_int64, c_void_p, c_void_p]),
    "tn_last_error": (ctypes.c_char_p, []),
    "tn_version": (ctypes.c_char_p, []),
}

EXPORTED = tuple(_SIGNATURES)

_lib = None


def load():
    """Load libtilenufft.so from the package tree (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"tilenufft native library not found at {LIB_PATH}; build it with "
            "`make -C deep-mr_amd/csrc` (or __graft_entry__.build())"
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, what="tilenufft call"):
    """Raise RuntimeError with the library's message for a non-zero status."""
    if rc != 0:
        msg = load().tn_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")
    return rc


def version():
    return load().tn_version().decode()
