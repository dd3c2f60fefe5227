"""Mesh voxelization through libspk.so (``spk_voxelize_mesh``).

``psacking_amd.voxelize_stl(path, resolution)`` is the reference's file entry
point (spectral_packing/bindings.cpp:285-302); ``voxelize_mesh`` applies the same
voxelization (packing.cpp:526-588 plus the tight crop of voxelGrid.cpp:166-178)
to an in-memory triangle mesh, e.g. one loaded by another mesh library.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def voxelize_mesh(positions, triangles, resolution: int = 128, device: int = 0) -> np.ndarray:
    """Surface voxels (int32 0/1, tight) of a triangle mesh.

    positions: (n, 3) float coordinates; triangles: (m, 3) vertex indices.
    Vertices at identical positions are merged first, as the reference does
    after loading an STL file.
    """
    lib = _lib.load()
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    tri = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1, 3)
    dims = (C.c_int32 * 3)()
    fp = pos.ctypes.data_as(C.POINTER(C.c_float))
    ip = tri.ctypes.data_as(C.POINTER(C.c_int32))

    def call(out, cap):
        rc = lib.spk_voxelize_mesh(device, fp, len(pos), ip, len(tri), int(resolution), dims, out, cap)
        if rc != 0:
            raise RuntimeError(f"voxelization failed: {_lib.last_error()}")

    call(None, 0)
    grid = np.empty(tuple(dims), dtype=np.int32)
    call(grid.ctypes.data_as(C.POINTER(C.c_int32)), grid.size)
    return grid
