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
    "spk_tray_get", "spk_tray_stats", "spk_tray_info", "spk_tray_clear", "spk_search", "spk_search_oriented",
    "spk_place", "spk_pack", "spk_pack_gather", "spk_pack_many", "spk_distance", "spk_correlate", "spk_collision_grid",
    "spk_interlock_mask", "spk_place_grid", "spk_bounds", "spk_save_vox", "spk_timing_enable",
    "spk_timing_get", "spk_timing_units", "spk_timing_reset", "spk_set_option", "spk_plan_info",
    "spk_tile_info",
    "spk_dev_alloc", "spk_dev_free", "spk_dev_copy", "spk_dev_gather",
    "spk_nn_create", "spk_nn_destroy", "spk_nn_num_params", "spk_nn_set_params",
    "spk_nn_get_params", "spk_nn_forward", "spk_nn_gradient", "spk_nn_train_step",
    "spk_nn_item_features", "spk_tray_features", "spk_pack_records", "spk_pack_learned",
    "spk_voxelize_mesh", "spk_voxelize_stl", "spk_feasible", "spk_distance_noalloc",
)


class SpkOrient(C.Structure):
    _fields_ = [("perm", C.c_int32 * 3), ("flip", C.c_int32 * 3)]


class SpkResult(C.Structure):
    _fields_ = [("pos", C.c_int32 * 3), ("found", C.c_int32), ("orient", C.c_int32),
                ("path", C.c_int32), ("score", C.c_double), ("collision", C.c_int64),
                ("proximity", C.c_int64), ("residual", C.c_float), ("reserved", C.c_int32)]


class SpkPackArgs(C.Structure):
    _fields_ = [("items", C.POINTER(C.c_int32)), ("dims", C.POINTER(C.c_int32)),
                ("n_items", C.c_int), ("order", C.POINTER(C.c_int32)),
                ("orients", C.POINTER(SpkOrient)), ("n_orient", C.c_int), ("penalty", C.c_int),
                ("interlocking_free", C.c_int), ("id_offset", C.c_int32),
                ("results", C.POINTER(SpkResult)),
                ("item_ptrs", C.POINTER(C.POINTER(C.c_int32)))]


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python psacking_amd/build.py`")
    lib = C.CDLL(LIB_PATH)
    lib.spk_last_error.restype = C.c_char_p
    vp, i32, i64, f32 = C.c_void_p, C.c_int, C.c_int64, C.c_float
    fp, dp = C.POINTER(C.c_float), C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    lib.spk_dev_alloc.argtypes = [i32, i64, C.POINTER(vp)]
    lib.spk_dev_free.argtypes = [i32, vp]
    lib.spk_dev_copy.argtypes = [i32, vp, vp, i64, i32]
    lib.spk_dev_gather.argtypes = [i32, vp, vp, i64, i64, C.POINTER(C.c_int64), i32]
    lib.spk_nn_create.argtypes = [i32, i32, C.POINTER(vp)]
    lib.spk_nn_destroy.argtypes = [vp]
    lib.spk_nn_num_params.argtypes = [vp, C.POINTER(i64)]
    lib.spk_nn_set_params.argtypes = [vp, fp, i64]
    lib.spk_nn_get_params.argtypes = [vp, fp, i64]
    lib.spk_nn_forward.argtypes = [vp, vp, vp, i32, fp]
    lib.spk_nn_gradient.argtypes = [vp, vp, vp, vp, vp, i32, fp, fp]
    lib.spk_nn_train_step.argtypes = [vp, vp, vp, vp, vp, i32, f32, fp]
    lib.spk_nn_item_features.argtypes = [vp, ip, i32, i32, i32, C.POINTER(SpkOrient), vp]
    lib.spk_tray_features.argtypes = [vp, vp]
    lib.spk_pack_records.argtypes = [vp, C.POINTER(SpkPackArgs), vp, dp]
    lib.spk_pack_learned.argtypes = [vp, vp, C.POINTER(SpkPackArgs), i32, f32, ip, ip]
    lib.spk_voxelize_mesh.argtypes = [i32, fp, i64, ip, i64, i32, ip, ip, i64]
    lib.spk_voxelize_stl.argtypes = [i32, C.c_char_p, i32, ip, ip, i64]
    lib.spk_feasible.argtypes = [vp, ip, i32, i32, i32, C.POINTER(SpkOrient), i32, ip, ip, fp]
    for name in ("spk_distance", "spk_distance_noalloc"):
        getattr(lib, name).argtypes = [vp, ip, i32, i32, i32, i32, i32, ip]
    _lib = lib
    return lib


def last_error() -> str:
    return load().spk_last_error().decode(errors="replace")
