"""ctypes view of libspk.so (the C-ABI in include/spk.h).

Used by the ABI-level tests and by bench.py; the Python API goes through the
pybind11 module ``_core`` instead.  Loading fails loudly when the library has
not been built.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspk.so")

# every entry point include/spk.h declares
EXPORTS = (
    "spk_last_error", "spk_version", "spk_ctx_create", "spk_ctx_destroy", "spk_tray_set",
    "spk_tray_get", "spk_tray_info", "spk_tray_clear", "spk_search", "spk_search_oriented",
    "spk_place", "spk_pack", "spk_pack_many", "spk_distance", "spk_correlate", "spk_collision_grid",
    "spk_interlock_mask", "spk_place_grid", "spk_bounds", "spk_save_vox", "spk_timing_enable",
    "spk_timing_get", "spk_timing_reset", "spk_set_option", "spk_plan_info",
)


class SpkOrient(C.Structure):
    _fields_ = [("perm", C.c_int32 * 3), ("flip", C.c_int32 * 3)]


class SpkResult(C.Structure):
    _fields_ = [("pos", C.c_int32 * 3), ("found", C.c_int32), ("orient", C.c_int32),
                ("path", C.c_int32), ("score", C.c_double), ("collision", C.c_int64),
                ("proximity", C.c_int64), ("residual", C.c_float), ("reserved", C.c_int32)]


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m psacking_amd.build`")
    lib = C.CDLL(LIB_PATH)
    lib.spk_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def last_error() -> str:
    return load().spk_last_error().decode(errors="replace")
